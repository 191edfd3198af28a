// dare::Matrix::{MatrixBlock, Terms, Trilinos, TrilinosSolver} over the MI355X C-ABI.
//
// Two Build paths, as in MatrixSystem/Trilinos.inl:50-60 / 287-385:
//  * Terms (device): the operator mb = ddt(phi); mb -= div(D grad(mb)); mb += div(tvd phi);
//    rhs += V s; boundary edits — evaluated per cell by the HIP assembly kernels with the
//    reference's operation order (dare_csr_build_structure / dare_csr_fill_values).
//  * any functor on a host MatrixBlock (Get/Set/Add/Remove/Finalize/SetCoefficient), rows
//    sorted like Tpetra's fillComplete, then uploaded (dare_csr_upload_host).
// The matrix (GetA), x and B live in HBM; TrilinosSolver::Solve(BumbleBee, "BICGSTAB2", ...)
// runs the device BiCGStab(2).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "Data.h"
#include "Stencils.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace dare {

// The subset of Teuchos::ParameterList the solver reads (BiCGStab2.inl:95-111).
class ParameterList {
public:
    using Value = std::variant<double, int, std::string>;
    template <typename T>
    void set(const std::string& k, T v) {
        if constexpr (std::is_floating_point_v<T>) values_[k] = static_cast<double>(v);
        else if constexpr (std::is_integral_v<T>) values_[k] = static_cast<int>(v);
        else values_[k] = std::string(v);
    }
    template <typename T>
    T get(const std::string& k, T dflt) const {
        auto it = values_.find(k);
        if (it == values_.end()) return dflt;
        if (auto p = std::get_if<T>(&it->second)) return *p;
        if constexpr (std::is_floating_point_v<T>)
            if (auto q = std::get_if<int>(&it->second)) return static_cast<T>(*q);
        throw std::runtime_error("parameter '" + k + "' has another type");
    }
    bool isParameter(const std::string& k) const { return values_.count(k) != 0; }

private:
    std::map<std::string, Value> values_;
};

}  // namespace dare

namespace dare::Matrix {

using CartesianNeighbor = dare::Grid::CartesianNeighbor;

// ------------------------------------------------------------------------------ MatrixBlock
// Host row builder for generic functors. MatrixBlockBase.inl:72-295 + MatrixBlock_Cartesian.inl:
// 128-593: 2*Dim+1 neighbour coefficients with a set-bitset per component, explicit
// (column, value) entries, and Finalize emitting W,S,B,C,T,N,E for the set neighbours.
template <typename Grid, std::size_t N>
class MatrixBlock {
public:
    using GridRepresentation = typename Grid::Representation;
    using Index = typename Grid::Index;
    using O = int64_t;
    using SC = double;
    static constexpr std::size_t Dim = Grid::DIMENSION;

    MatrixBlock() = default;
    void Initialize(const GridRepresentation* g, O node) {
        g_ = g;
        node_ = node;
        ind_internal_ = g->MapOrdinalToIndexLocalInternal(static_cast<int32_t>(node));
        for (std::size_t n = 0; n < N; n++) {
            nb_[n].fill(0.);
            bits_[n] = 0;
            ord_[n].clear();
            coef_[n].clear();
            rhs_[n] = 0.;
            guess_[n] = 0.;
        }
    }
    const GridRepresentation* GetRepresentation() const { return g_; }
    O GetNode() const { return node_; }
    O GetRow(std::size_t n) const { return node_ * static_cast<O>(N) + static_cast<O>(n); }
    int32_t GetLocalOrdinal() const { return static_cast<int32_t>(node_); }
    static constexpr bool IsGlobal() { return true; }
    constexpr std::size_t GetNumComponents() const { return N; }
    Index GetIndexInternal() const { return ind_internal_; }
    Index GetIndex() const { return g_->MapInternalToLocal(ind_internal_); }

    SC& Get(std::size_t n, CartesianNeighbor cnb) {
        const int p = static_cast<int>(cnb);
        bits_[n] |= static_cast<uint8_t>(1u << p);
        return nb_[n][p];
    }
    SC Get(std::size_t n, CartesianNeighbor cnb) const { return nb_[n][static_cast<int>(cnb)]; }
    template <CartesianNeighbor CNB>
    SC& Get(std::size_t n) { return Get(n, CNB); }
    void Remove(std::size_t n, CartesianNeighbor cnb) {
        bits_[n] &= static_cast<uint8_t>(~(1u << static_cast<int>(cnb)));
    }
    template <CartesianNeighbor CNB>
    void Remove(std::size_t n) { Remove(n, CNB); }
    bool IsSet(std::size_t n, CartesianNeighbor cnb) const { return bits_[n] & (1u << static_cast<int>(cnb)); }

    SC& GetRhs(std::size_t n) { return rhs_[n]; }
    SC GetRhs(std::size_t n) const { return rhs_[n]; }
    void SetRhs(std::size_t n, SC v) { rhs_[n] = v; }
    SC& GetInitialGuess(std::size_t n) { return guess_[n]; }
    SC GetInitialGuess(std::size_t n) const { return guess_[n]; }
    void SetInitialGuess(std::size_t n, SC v) { guess_[n] = v; }

    // ---- stencils (MatrixBlock_Cartesian.inl:487-593): every position of the stencil is
    // written through Get (which marks it set), centre first, then W, E, S, N, B, T
    using CenterStencil = Data::CenterMatrixStencil<Grid, SC, N>;
    using CenterSlice = Data::CenterMatrixStencil<Grid, SC, 1>;
    MatrixBlock& Set(const CenterStencil& s) {
        for (std::size_t n = 0; n < N; n++) {
            for (std::size_t p = 0; p < 2 * Dim + 1; p++) Get(n, static_cast<CartesianNeighbor>(p)) = s.GetData()[n][p];
            GetRhs(n) = s.GetRHS()[n];
        }
        return *this;
    }
    MatrixBlock& Set(std::size_t n, const CenterSlice& s) {
        for (std::size_t p = 0; p < 2 * Dim + 1; p++) Get(n, static_cast<CartesianNeighbor>(p)) = s.GetData()[0][p];
        GetRhs(n) = s.GetRHS()[0];
        return *this;
    }
    MatrixBlock& Add(const CenterStencil& s) {
        for (std::size_t n = 0; n < N; n++) {
            for (std::size_t p = 0; p < 2 * Dim + 1; p++) Get(n, static_cast<CartesianNeighbor>(p)) += s.GetData()[n][p];
            GetRhs(n) += s.GetRHS()[n];
        }
        return *this;
    }
    MatrixBlock& Add(std::size_t n, const CenterSlice& s) {
        for (std::size_t p = 0; p < 2 * Dim + 1; p++) Get(n, static_cast<CartesianNeighbor>(p)) += s.GetData()[0][p];
        GetRhs(n) += s.GetRHS()[0];
        return *this;
    }
    // Subtract(s) = Add(-1 * s): coefficients and right-hand side negated first (:576-593)
    MatrixBlock& Subtract(const CenterStencil& s) { return Add(-1. * s); }
    MatrixBlock& Subtract(std::size_t n, const CenterSlice& s) { return Add(n, -1. * s); }
    MatrixBlock& operator=(const CenterStencil& s) { return Set(s); }
    MatrixBlock& operator+=(const CenterStencil& s) { return Add(s); }
    MatrixBlock& operator-=(const CenterStencil& s) { return Add(-1. * s); }

    // MatrixBlockBase: explicit entries
    void Resize(std::size_t n, std::size_t size) {
        const std::size_t old = ord_[n].size();
        ord_[n].resize(size);
        coef_[n].resize(size);
        for (std::size_t i = old; i < size; i++) {
            ord_[n][i] = -1;
            coef_[n][i] = 0.;
        }
    }
    std::size_t GetNumEntries(std::size_t n) const {
        std::size_t k = 0;
        for (auto o : ord_[n]) k += o > -1;
        return k;
    }
    void SetCoefficient(std::size_t n, O col, SC v) {      // MatrixBlockBase.inl:229-254
        for (std::size_t i = 0; i < ord_[n].size(); i++) {
            if (ord_[n][i] == col || ord_[n][i] == -1) {
                ord_[n][i] = col;
                coef_[n][i] = v;
                return;
            }
        }
    }
    SC& GetCoefficientByOrdinal(std::size_t n, O col) {    // MatrixBlockBase.inl:198-215
        for (std::size_t i = 0; i < ord_[n].size(); i++)
            if (ord_[n][i] == col) return coef_[n][i];
        ord_[n].push_back(col);
        coef_[n].push_back(0.);
        return coef_[n].back();
    }
    const std::vector<O>& GetColumnOrdinals(std::size_t n) const { return ord_[n]; }
    const std::vector<SC>& GetColumnValues(std::size_t n) const { return coef_[n]; }

    void Finalize() {                                       // MatrixBlock_Cartesian.inl:128-240
        static const int order[7] = {1, 3, 5, 0, 6, 4, 2};
        for (std::size_t n = 0; n < N; n++) {
            O cols[7];
            SC vals[7];
            std::size_t k = 0;
            for (int q = 0; q < 7; q++) {
                const int p = order[q];
                if (p > 0 && (p + 1) / 2 > static_cast<int>(Dim)) continue;
                if (!(bits_[n] & (1u << p))) continue;
                Index nb = ind_internal_;
                if (p > 0) nb[(p - 1) / 2] += (p % 2 == 1) ? -1 : 1;
                cols[k] = g_->MapIndexToOrdinalGlobalInternal(nb, true) * static_cast<O>(N) + static_cast<O>(n);
                vals[k] = nb_[n][p];
                k++;
            }
            SetCoefficients(n, cols, vals, k);
        }
    }

    // the row as Tpetra's fillComplete leaves it: ascending columns, equal columns summed
    void SortedRow(std::size_t n, std::vector<std::pair<O, SC>>* out) const {
        out->clear();
        for (std::size_t i = 0; i < ord_[n].size(); i++)
            if (ord_[n][i] > -1) out->push_back({ord_[n][i], coef_[n][i]});
        std::stable_sort(out->begin(), out->end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        std::size_t w = 0;
        for (std::size_t i = 0; i < out->size(); i++) {
            if (w > 0 && (*out)[w - 1].first == (*out)[i].first) (*out)[w - 1].second += (*out)[i].second;
            else (*out)[w++] = (*out)[i];
        }
        out->resize(w);
    }

private:
    void SetCoefficients(std::size_t n, const O* cols, const SC* vals, std::size_t sz) {   // :256-295
        Resize(n, GetNumEntries(n) + sz);
        const bool empty = ord_[n].empty() || ord_[n][0] < 0;
        if (empty) {
            for (std::size_t i = 0; i < sz; i++) {
                ord_[n][i] = cols[i];
                coef_[n][i] = vals[i];
            }
        } else {
            for (std::size_t i = 0; i < sz; i++) SetCoefficient(n, cols[i], vals[i]);
        }
        Resize(n, GetNumEntries(n));
    }

    const GridRepresentation* g_ = nullptr;
    O node_ = 0;
    Index ind_internal_{};
    std::array<std::array<SC, 7>, N> nb_{};
    std::array<uint8_t, N> bits_{};
    std::array<std::vector<O>, N> ord_;
    std::array<std::vector<SC>, N> coef_;
    std::array<SC, N> rhs_{}, guess_{};
};

// ------------------------------------------------------------------------------ device objects
class DeviceCsr {
public:
    explicit DeviceCsr(std::shared_ptr<Device> dev) : dev_(std::move(dev)) {
        Check(dare_csr_create(dev_->Get(), &h_), "dare_csr_create");
    }
    ~DeviceCsr() { dare_csr_destroy(h_); }
    DeviceCsr(const DeviceCsr&) = delete;
    DeviceCsr& operator=(const DeviceCsr&) = delete;
    dare_csr* Get() const { return h_; }
    int64_t getLocalNumRows() const {
        int64_t nr = 0;
        Check(dare_csr_info(h_, &nr, nullptr, nullptr, nullptr, nullptr), "dare_csr_info");
        return nr;
    }
    int64_t getLocalNumEntries() const {
        int64_t nnz = 0;
        Check(dare_csr_info(h_, nullptr, &nnz, nullptr, nullptr, nullptr), "dare_csr_info");
        return nnz;
    }
    // host copy of the CSR (Tpetra getLocalRowCopy for all rows)
    void Download(std::vector<int32_t>* rp, std::vector<int32_t>* col, std::vector<double>* val) const {
        const int64_t nr = getLocalNumRows(), nnz = getLocalNumEntries();
        rp->resize(nr + 1);
        col->resize(nnz);
        val->resize(nnz);
        Check(dare_csr_download(dev_->Get(), h_, rp->data(), col->data(), val->data()), "dare_csr_download");
    }
    // A.apply(x, y) (BiCGStab2.inl A21)
    void apply(const double* x, double* y) const { Check(dare_spmv(dev_->Get(), h_, x, y), "dare_spmv"); }
    Device& GetDevice() const { return *dev_; }

private:
    std::shared_ptr<Device> dev_;
    dare_csr* h_ = nullptr;
};

class DeviceVector {      // a Tpetra::Vector in HBM
public:
    DeviceVector(std::shared_ptr<Device> dev, std::size_t n) : buf_(std::move(dev), n) {}
    double* data() { return buf_.data(); }
    const double* data() const { return buf_.data(); }
    std::size_t getLocalLength() const { return buf_.size(); }
    std::vector<double> getData() const {
        std::vector<double> h(buf_.size());
        buf_.Download(h.data());
        return h;
    }
    void putScalar(double v) { Check(dare_fill_f64(buf_.GetDevice().Get(), buf_.data(), buf_.size(), v), "fill"); }
    double dot(const DeviceVector& o) const {
        double d = 0.;
        Check(dare_dot(buf_.GetDevice().Get(), buf_.size(), buf_.data(), o.data(), &d), "dare_dot");
        return d;
    }
    double norm2() const { return std::sqrt(dot(*this)); }
    // this = beta * this + alpha * x (Tpetra update)
    void update(double alpha, const DeviceVector& x, double beta) {
        Check(dare_axpby(buf_.GetDevice().Get(), buf_.size(), alpha, x.data(), beta, buf_.data()), "dare_axpby");
    }
    DeviceBuffer<double>& Buffer() { return buf_; }

private:
    DeviceBuffer<double> buf_;
};

// The device operator (the functor of the fast path). Face ids: 0 W, 1 E, 2 S, 3 N, 4 B, 5 T.
struct Terms {
    dare_terms t{};
    std::shared_ptr<void> keep;   // owner of device arrays referenced by t
    Terms() { t.conv_limiter = -1; }
    Terms& DDT(double dt) { t.dt = dt; return *this; }
    Terms& Diffusion(double D) { t.diff_const = D; return *this; }
    Terms& DiffusionPerComponent(const std::vector<double>& D) {
        t.use_diff_comp = 1;
        for (std::size_t c = 0; c < D.size() && c < 8; c++) t.diff_comp[c] = D[c];
        return *this;
    }
    Terms& DiffusionField(const double* dev_cell_D) { t.diff_cell = dev_cell_D; return *this; }
    Terms& TVD(int limiter, const double* ux, const double* uy, const double* uz) {
        t.conv_limiter = limiter;
        t.vel[0] = ux;
        t.vel[1] = uy;
        t.vel[2] = uz;
        return *this;
    }
    Terms& Source(const double* dev_cell_s) { t.source = dev_cell_s; return *this; }
    Terms& Reaction(const double* dev_blocks) { t.react = dev_blocks; return *this; }
    Terms& Dirichlet(int face, double v) { t.bc_type[face] = DARE_BC_DIRICHLET; t.bc_value[face] = v; return *this; }
    Terms& Neumann(int face) { t.bc_type[face] = DARE_BC_NEUMANN; return *this; }
    // the boundary cells of this face carry v themselves (identity rows): Dirichlet on a
    // staggered axis, whose first / last cells sit on the boundary faces
    Terms& Fixed(int face, double v) { t.bc_type[face] = DARE_BC_FIXED; t.bc_value[face] = v; return *this; }
    // the time scheme of the diffusion and convection divergences, as Divergence<Grid, Scheme>
    // (EULER_BACKWARD, EULER_FORWARD, CRANK_NICHOLSON, ADAMS_BASHFORT, ADAMS_MOULTON)
    template <typename Scheme>
    Terms& DivergenceScheme() {
        const auto w = Scheme::template GetWeights<double>();
        t.div_levels = static_cast<int32_t>(w.size() < 3 ? w.size() : 3);
        for (std::size_t l = 0; l < w.size() && l < 3; l++) t.div_weights[l] = w[l];
        return *this;
    }
};

// A built preconditioner (the RCP<OperatorType> M of the reference): which right
// preconditioner the device solver applies (dare_solver_params.precond)
struct Preconditioner {
    int32_t precond;
};

// ------------------------------------------------------------------------------ Trilinos
template <typename SC>
class Trilinos {
public:
    using MatrixType = DeviceCsr;
    using VecType = DeviceVector;

    explicit Trilinos(std::shared_ptr<Device> dev = nullptr) : dev_(std::move(dev)) {}

    // device fast path (the functor is `terms`); structure built once, values re-filled
    // in place on later calls (BuildReplace semantics, Trilinos.inl:389-457)
    template <typename Grid, std::size_t N>
    void Build(const typename Grid::Representation& grep, Data::GridVector<Grid, SC, N>& field,
               const Terms& terms, bool rebuild = false) {
        const dare_grid g = field.Descriptor();
        const std::size_t nrows = static_cast<std::size_t>(grep.GetNumberLocalCellsInternal()) * N;
        if (!A_ || !structured_ || rebuild || nrows != x_->getLocalLength()) {
            A_ = std::make_shared<DeviceCsr>(dev());
            Check(dare_csr_build_structure(dev()->Get(), A_->Get(), &g, &terms.t), "dare_csr_build_structure");
            x_ = std::make_shared<DeviceVector>(dev(), nrows);
            B_ = std::make_shared<DeviceVector>(dev(), nrows);
            structured_ = true;
        }
        Check(dare_csr_fill_values(dev()->Get(), A_->Get(), &g, &terms.t, field.GetDeviceView(), B_->data(),
                                   x_->data()),
              "dare_csr_fill_values");
    }

    // The rows of a generic functor, assembled on the host: per-cell MatrixBlock, Finalize, rows
    // sorted with equal columns summed (what fillComplete leaves), x0 from the field
    // (Trilinos.inl:287-385). The functor runs concurrently on OpenMP threads, as in the
    // reference (Trilinos.inl:308), so it must be thread-safe; each thread owns a contiguous
    // range of cells and the rows are joined in cell order, so the result does not depend on
    // the thread count.
    struct HostRows {
        std::vector<int32_t> row_ptr, col;
        std::vector<double> val, rhs, x0;
    };
    template <typename Grid, std::size_t N, typename Lambda>
    static HostRows AssembleHost(const typename Grid::Representation& grep, Data::GridVector<Grid, SC, N>& field,
                                 Lambda functor) {
        field.SyncToHost();
        const int32_t ncell = grep.GetNumberLocalCellsInternal();
        const int64_t nrows = static_cast<int64_t>(ncell) * N;
        int nth = 1;
#ifdef _OPENMP
        nth = omp_get_max_threads();
        if (nth > ncell) nth = ncell > 0 ? ncell : 1;
#endif
        struct Part {
            std::vector<int32_t> len, col;
            std::vector<double> val;
            std::string error;
        };
        std::vector<Part> parts(nth);
        HostRows out;
        out.rhs.resize(nrows);
        out.x0.resize(nrows);
#pragma omp parallel num_threads(nth)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            Part& P = parts[t];
            const int32_t lo = static_cast<int32_t>(static_cast<int64_t>(ncell) * t / nth);
            const int32_t hi = static_cast<int32_t>(static_cast<int64_t>(ncell) * (t + 1) / nth);
            try {
                MatrixBlock<Grid, N> mb;
                std::vector<std::pair<int64_t, double>> row;
                for (int32_t node = lo; node < hi; node++) {
                    mb.Initialize(&grep, node);
                    const int32_t full = grep.MapInternalToLocal(node);
                    for (std::size_t n = 0; n < N; n++) mb.SetInitialGuess(n, static_cast<const Data::GridVector<Grid, SC, N>&>(field).At(full, n));
                    functor(&mb);
                    mb.Finalize();
                    for (std::size_t n = 0; n < N; n++) {
                        const std::size_t r = static_cast<std::size_t>(node) * N + n;
                        mb.SortedRow(n, &row);
                        for (const auto& e : row) {
                            if (e.first < 0 || e.first >= nrows)
                                throw std::runtime_error("functor produced a column outside the matrix");
                            P.col.push_back(static_cast<int32_t>(e.first));
                            P.val.push_back(e.second);
                        }
                        P.len.push_back(static_cast<int32_t>(row.size()));
                        out.rhs[r] = mb.GetRhs(n);
                        out.x0[r] = mb.GetInitialGuess(n);
                    }
                }
            } catch (const std::exception& e) {
                P.error = e.what();
            }
        }
        std::size_t nnz = 0;
        for (const Part& P : parts) {
            if (!P.error.empty()) throw std::runtime_error(P.error);
            nnz += P.col.size();
        }
        if (nnz >= (std::size_t(1) << 31)) throw std::runtime_error("nnz exceeds 32-bit row pointers");
        out.row_ptr.reserve(nrows + 1);
        out.row_ptr.push_back(0);
        out.col.reserve(nnz);
        out.val.reserve(nnz);
        for (const Part& P : parts) {
            for (int32_t l : P.len) out.row_ptr.push_back(out.row_ptr.back() + l);
            out.col.insert(out.col.end(), P.col.begin(), P.col.end());
            out.val.insert(out.val.end(), P.val.begin(), P.val.end());
        }
        return out;
    }

    // generic path: any functor on a host MatrixBlock, uploaded to the device (Trilinos.inl:287-385)
    template <typename Grid, std::size_t N, typename Lambda>
    void Build(const typename Grid::Representation& grep, Data::GridVector<Grid, SC, N>& field, Lambda functor,
               bool = false) {
        HostRows h = AssembleHost(grep, field, functor);
        const std::size_t nrows = h.rhs.size();
        A_ = std::make_shared<DeviceCsr>(dev());
        Check(dare_csr_upload_host(dev()->Get(), A_->Get(), static_cast<int64_t>(nrows),
                                   static_cast<int64_t>(h.col.size()), h.row_ptr.data(), h.col.data(), h.val.data()),
              "dare_csr_upload_host");
        x_ = std::make_shared<DeviceVector>(dev(), nrows);
        B_ = std::make_shared<DeviceVector>(dev(), nrows);
        x_->Buffer().Upload(h.x0.data());
        B_->Buffer().Upload(h.rhs.data());
        structured_ = false;
    }

    std::shared_ptr<MatrixType>& GetA() { return A_; }
    std::shared_ptr<VecType>& GetX() { return x_; }
    std::shared_ptr<VecType>& GetB() { return B_; }
    std::shared_ptr<Preconditioner>& GetM() { return M_; }   // null, as Trilinos::GetM before one is set

    // Trilinos.inl:168-177: x[node*N + n] -> field.At(MapInternalToLocal(node), n), on the device
    template <typename Grid, std::size_t N>
    void CopyTo(Data::GridVector<Grid, SC, N>* gvec) {
        const dare_grid g = gvec->Descriptor();
        Check(dare_copy_to_field(dev()->Get(), &g, x_->data(), gvec->GetDeviceView()), "dare_copy_to_field");
        dev()->Synchronize();
        gvec->SyncToHost();
    }

private:
    const std::shared_ptr<Device>& dev() {
        if (!dev_) dev_ = Device::Default();
        return dev_;
    }
    std::shared_ptr<Device> dev_;
    std::shared_ptr<MatrixType> A_;
    std::shared_ptr<VecType> x_, B_;
    std::shared_ptr<Preconditioner> M_;
    bool structured_ = false;
};

// ------------------------------------------------------------------------------ TrilinosSolver
enum class SolverPackage { Amesos2, Belos, BumbleBee };      // TrilinosSolver.h:60-64
enum class PreCondPackage { None, Ifpack2, MueLu };          // TrilinosSolver.h:55-59
enum class ReturnType { Converged, Unconverged };            // Belos::ReturnType

template <typename SC>
class TrilinosSolver {
public:
    // TrilinosSolver.inl:71-93 + CreateSolver(BumbleBee, "BICGSTAB2") (:150-152)
    ReturnType Solve(SolverPackage pack, const std::string& type, const std::shared_ptr<Preconditioner>& M,
                     const std::shared_ptr<DeviceCsr>& A, const std::shared_ptr<DeviceVector>& x,
                     const std::shared_ptr<DeviceVector>& B, const std::shared_ptr<ParameterList>& params) {
        if (pack != SolverPackage::BumbleBee || type != "BICGSTAB2")
            throw std::runtime_error("this library provides BumbleBee/BICGSTAB2 only (got '" + type + "')");
        dare_solver_params p{};
        p.precond = M ? M->precond : 0;      // right preconditioner (BiCGStab2.inl:353-365)
        p.tol = params ? params->get<double>("Convergence Tolerance", 1e-8) : 1e-8;   // BiCGStab2.inl:29-30
        p.max_it = params ? params->get<int>("Maximum Iterations", 100) : 100;
        p.tol_kind = 0;                                                               // RNORM (:34)
        dare_solver_result r{};
        Check(dare_bicgstab2_solve(A->GetDevice().Get(), A->Get(), B->data(), x->data(), &p, &r),
              "dare_bicgstab2_solve");
        num_iter_ = r.iterations;
        residual_ = r.residual;
        return r.converged ? ReturnType::Converged : ReturnType::Unconverged;
    }
    ReturnType Solve(SolverPackage pack, const std::string& type, const std::shared_ptr<DeviceCsr>& A,
                     const std::shared_ptr<DeviceVector>& x, const std::shared_ptr<DeviceVector>& B,
                     const std::shared_ptr<ParameterList>& params) {
        return Solve(pack, type, nullptr, A, x, B, params);
    }
    // TrilinosSolver.inl:97-127 + CreatePreconditionerIfPack2 (:164-175) and
    // CreatePreconditionerMueLu (:180-246). This library provides
    //  - Ifpack2 "RELAXATION" of type "Jacobi" with one sweep and damping 1, i.e.
    //    M^-1 = diag(A)^-1 (any operator);
    //  - MueLu "AMG": the aggregation multigrid V-cycle of the device solver (2x2x2 aggregates,
    //    damped-Jacobi smoothing), for device-assembled one-component operators on a 3-D
    //    Cartesian grid (Build(grep, field, terms)); MueLu's parameter list ("problem: type",
    //    "verbosity") is accepted and not used, and the solve refuses operators the hierarchy
    //    cannot coarsen (generic CSR from a host functor, two components).
    // Both are applied on the right; None returns null as the reference does. Other types throw
    // (the reference prints an error and terminates).
    std::shared_ptr<Preconditioner> BuildPreconditioner(PreCondPackage pack, const std::string& type,
                                                        const std::shared_ptr<ParameterList>& param,
                                                        const std::shared_ptr<DeviceCsr>&) {
        if (pack == PreCondPackage::None) return nullptr;
        if (pack == PreCondPackage::MueLu && type == "AMG")
            return std::make_shared<Preconditioner>(Preconditioner{DARE_PRECOND_MG});
        const std::string relax = param ? param->get<std::string>("relaxation: type", "Jacobi") : "Jacobi";
        const int sweeps = param ? param->get<int>("relaxation: sweeps", 1) : 1;
        const double omega = param ? param->get<double>("relaxation: damping factor", 1.) : 1.;
        if (pack == PreCondPackage::Ifpack2 && type == "RELAXATION" && relax == "Jacobi" && sweeps == 1 && omega == 1.)
            return std::make_shared<Preconditioner>(Preconditioner{DARE_PRECOND_JACOBI});
        throw std::runtime_error("Preconditioner of type " + type +
                                 " is unknown to this library (Ifpack2 RELAXATION / Jacobi, one sweep, damping 1;"
                                 " MueLu AMG)");
    }
    int GetNumIterations() const { return num_iter_; }      // TrilinosSolver.inl:263-265
    double GetResidual() const { return residual_; }

private:
    int num_iter_ = 0;
    double residual_ = 0.;
};

}  // namespace dare::Matrix
