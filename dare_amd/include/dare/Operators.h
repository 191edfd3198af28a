// The per-cell operator classes of the reference's call surface on the host mirror:
//   dare::Matrix::Gradient<Grid>            Grid/Cartesian/Gradient_Cartesian.inl:28-186
//   dare::Matrix::Divergence<Grid, Scheme>  Grid/Cartesian/Divergence_Cartesian.inl:30-235,362-378
//   dare::Matrix::TVD<Grid, SC, Limiter>    Grid/Cartesian/TVD_Cartesian.inl:27-261
//   dare::Matrix::DDT<Grid, Scheme>         Equations/DDT.inl:28-109
//   limiters CDS, UPWIND, VANALBADA, MINMOD (Equations/FluxLimiter.h:35-173) and VANLEER
//   time schemes (Equations/TimeDiscretizationSchemes.h:28-76, DDT.h:35-44)
//   dare::math::InterpolateToFace           Grid/Cartesian/Interpolation_Cartesian.h:59-354
// A functor written against the reference (`(*mb) = ddt(field); (*mb) -= div(grad(*mb));
// mb->Add(n, div(tvd * phi).GetSlice(n))`) evaluates here with the reference's floating-point
// operation order, so the host MatrixBlock path builds the same matrix as the device assembly
// of the same terms (dare_csr_fill_values) and as the CPU oracle, bit for bit.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <tuple>
#include <type_traits>

#include "Data.h"
#include "MatrixSystem.h"
#include "Stencils.h"

namespace dare::Matrix {

// ------------------------------------------------------------------------------ limiters
// Psi(r), element-wise on a Vector (FluxLimiter.h:35-173); NaN -> 0 for VANALBADA, MINMOD and
// VANLEER. VANLEER is not in the reference: (r + |r|) / (1 + |r|), 2 at r = +inf, 0 at -inf.
#define DARE_LIMITER_VECTOR_FORM                                                  \
    template <std::size_t N, typename SC>                                         \
    static utils::Vector<N, SC> GetValue(const utils::Vector<N, SC>& r) {        \
        utils::Vector<N, SC> ret;                                                 \
        for (std::size_t n = 0; n < N; n++) ret[n] = GetValue(r[n]);              \
        return ret;                                                               \
    }
struct CDS {
    DARE_LIMITER_VECTOR_FORM
    template <typename SC>
    static SC GetValue(const SC&) { return 1.; }
};
struct UPWIND {
    DARE_LIMITER_VECTOR_FORM
    template <typename SC>
    static SC GetValue(const SC&) { return 0.; }
};
struct VANALBADA {
    DARE_LIMITER_VECTOR_FORM
    template <typename SC>
    static SC GetValue(SC r) {
        r = std::isnan(r) ? SC(0) : r;
        const SC r2 = r * r;
        return (r + r2) / (1. + r2);
    }
};
struct MINMOD {
    DARE_LIMITER_VECTOR_FORM
    template <typename SC>
    static SC GetValue(SC r) {
        r = std::isnan(r) ? SC(0) : r;
        return std::max(SC(0), std::min(r, SC(1)));
    }
};
struct VANLEER {
    DARE_LIMITER_VECTOR_FORM
    template <typename SC>
    static SC GetValue(SC r) {
        if (std::isnan(r)) return 0.;
        if (std::isinf(r)) return r > 0 ? 2. : 0.;
        const SC a = std::fabs(r);
        return (r + a) / (1. + a);
    }
};
#undef DARE_LIMITER_VECTOR_FORM

// ------------------------------------------------------------------------------ time schemes
// weights of the levels n+1, n, n-1 (TimeDiscretizationSchemes.h:28-76)
struct EULER_BACKWARD {
    static const std::size_t NUM_TIMESTEPS{0};
    template <typename SC>
    static utils::Vector<1, SC> GetWeights() { return utils::Vector<1, SC>(SC(1.)); }
};
struct EULER_FORWARD {
    static const std::size_t NUM_TIMESTEPS{1};
    template <typename SC>
    static utils::Vector<2, SC> GetWeights() { return utils::Vector<2, SC>(SC(0.), SC(1.)); }
};
struct CRANK_NICHOLSON {
    static const std::size_t NUM_TIMESTEPS{1};
    template <typename SC>
    static utils::Vector<2, SC> GetWeights() { return utils::Vector<2, SC>(SC(0.5), SC(0.5)); }
};
struct ADAMS_BASHFORT {
    static const std::size_t NUM_TIMESTEPS{2};
    template <typename SC>
    static utils::Vector<3, SC> GetWeights() { return utils::Vector<3, SC>(SC(0.), SC(1.5), SC(-0.5)); }
};
struct ADAMS_MOULTON {
    static const std::size_t NUM_TIMESTEPS{2};
    template <typename SC>
    static utils::Vector<3, SC> GetWeights() {
        return utils::Vector<3, SC>(SC(5. / 12.), SC(8. / 12.), SC(-1. / 12.));
    }
};
namespace timeschemes {
struct IMPLICIT {      // DDT.h:35-44: phi^{n+1} and phi^n, both weight 1
    static const std::size_t NUM_TIMESTEPS{1};
    template <typename SC>
    static utils::Vector<2, SC> GetWeights() { return utils::Vector<2, SC>(SC(1.), SC(1.)); }
};
}  // namespace timeschemes

namespace detail {
template <typename T>
struct is_field : std::false_type {};
template <typename G, typename SC, std::size_t N>
struct is_field<Data::Field<G, SC, N>> : std::true_type {};
template <typename T>
struct is_gridvector : std::false_type {};
template <typename G, typename SC, std::size_t N>
struct is_gridvector<Data::GridVector<G, SC, N>> : std::true_type {};
template <typename T>
struct is_face_matrix : std::false_type {};
template <typename G, typename SC, std::size_t N>
struct is_face_matrix<Data::FaceMatrixStencil<G, SC, N>> : std::true_type {};
template <typename T>
struct is_face_value : std::false_type {};
template <typename G, typename SC, std::size_t N>
struct is_face_value<Data::FaceValueStencil<G, SC, N>> : std::true_type {};
}  // namespace detail

// ------------------------------------------------------------------------------ Gradient
template <typename Grid>
class Gradient {
public:
    static constexpr std::size_t Dim = Grid::DIMENSION;
    using GridType = Grid;
    using GridRepresentation = typename Grid::Representation;
    using LO = typename Grid::LocalOrdinalType;
    using Index = typename Grid::Index;
    using Positions = typename Grid::NeighborID;

    // dn_r = 1 / dx per axis (Gradient_Cartesian.inl:28-34)
    Gradient(const GridRepresentation& grid, LO ordinal_internal) : grid_(&grid), ordinal_internal_(ordinal_internal) {
        dn_r_.SetAllValues(1.);
        dn_r_ /= grid.GetDistances();
    }

    // implicit form: face W (nb -dn_r, c +dn_r), E (nb +dn_r, c -dn_r), ... (:39-58)
    template <std::size_t N>
    Data::FaceMatrixStencil<Grid, double, N> operator()(const MatrixBlock<Grid, N>&) const {
        Data::FaceMatrixStencil<Grid, double, N> s;
        s.SetValues(Positions::WEST, 0, -dn_r_[0], dn_r_[0]);
        s.SetValues(Positions::EAST, 0, dn_r_[0], -dn_r_[0]);
        if constexpr (Dim > 1) {
            s.SetValues(Positions::SOUTH, 0, -dn_r_[1], dn_r_[1]);
            s.SetValues(Positions::NORTH, 0, dn_r_[1], -dn_r_[1]);
        }
        if constexpr (Dim > 2) {
            s.SetValues(Positions::BOTTOM, 0, -dn_r_[2], dn_r_[2]);
            s.SetValues(Positions::TOP, 0, dn_r_[2], -dn_r_[2]);
        }
        for (std::size_t n = 1; n < N; n++) {
            s.GetDataCenter()[n] = s.GetDataCenter()[n - 1];
            s.GetDataNeighbor()[n] = s.GetDataNeighbor()[n - 1];
        }
        return s;
    }

    // face gradients of a field: W = (phiC - phiW) dn_r, E = (phiE - phiC) dn_r, ... (:62-96)
    template <typename SC, std::size_t N>
    Data::FaceValueStencil<Grid, SC, N> operator()(const Data::GridVector<Grid, SC, N>& field) const {
        Data::FaceValueStencil<Grid, SC, N> s;
        for (std::size_t n = 0; n < N; n++) {
            auto one = (*this)(field, n);
            for (std::size_t f = 0; f < 2 * Dim; f++) s.GetData()[n][f] = one.GetData()[0][f];
        }
        return s;
    }
    template <typename SC, std::size_t N>
    Data::FaceValueStencil<Grid, SC, 1> operator()(const Data::GridVector<Grid, SC, N>& field, std::size_t n) const {
        Data::FaceValueStencil<Grid, SC, 1> s;
        const Index ind = grid_->MapInternalToLocal(grid_->MapOrdinalToIndexLocalInternal(ordinal_internal_));
        for (std::size_t d = 0; d < Dim; d++) {
            Index lo = ind, hi = ind;
            lo[d] -= 1;
            hi[d] += 1;
            const SC c = field.At(ind, n);
            s.SetValue(::dare::Grid::ToCartesianNeighbor(2 * d + 1), 0, (c - field.At(lo, n)) * dn_r_[d]);
            s.SetValue(::dare::Grid::ToCartesianNeighbor(2 * d + 2), 0, (field.At(hi, n) - c) * dn_r_[d]);
        }
        return s;
    }
    // the same from the values of a CenterValueStencil (:100-131)
    template <typename SC, std::size_t N>
    Data::FaceValueStencil<Grid, SC, N> operator()(const Data::CenterValueStencil<Grid, SC, N>& s_c) const {
        Data::FaceValueStencil<Grid, SC, N> s;
        for (std::size_t n = 0; n < N; n++)
            for (std::size_t d = 0; d < Dim; d++) {
                const auto lo = ::dare::Grid::ToCartesianNeighbor(2 * d + 1), hi = ::dare::Grid::ToCartesianNeighbor(2 * d + 2);
                s.SetValue(lo, n, (s_c.GetValue(Positions::CENTER, n) - s_c.GetValue(lo, n)) * dn_r_[d]);
                s.SetValue(hi, n, (s_c.GetValue(hi, n) - s_c.GetValue(Positions::CENTER, n)) * dn_r_[d]);
            }
        return s;
    }

private:
    const GridRepresentation* grid_;
    LO ordinal_internal_;
    typename Grid::VecSC dn_r_;
};

// ------------------------------------------------------------------------------ Divergence
template <typename Grid, typename TimeDiscretization = EULER_BACKWARD>
class Divergence {
public:
    static constexpr std::size_t Dim = Grid::DIMENSION;
    using GridType = Grid;
    using GridRepresentation = typename Grid::Representation;
    using LO = typename Grid::LocalOrdinalType;
    using Index = typename Grid::Index;
    using Positions = typename Grid::NeighborID;
    static const std::size_t NUM_TIMESTEPS = TimeDiscretization::NUM_TIMESTEPS;
    static const std::size_t NUM_TFIELDS = NUM_TIMESTEPS + 1;

    Divergence(const GridRepresentation& grid, LO ordinal_internal)
        : A_(grid.GetFaceArea()),
          ind_(grid.MapInternalToLocal(grid.MapOrdinalToIndexLocalInternal(ordinal_internal))),
          grep_(&grid) {}

    // div(a1, ..., ak, s): s a FaceMatrixStencil or Field (matrix form, -> CenterMatrixStencil) or
    // a FaceValueStencil (value form, -> Vector<N>), multiplied first by a1..ak (scalars or
    // FaceValueStencils) (Divergence_Cartesian.inl:39-88)
    template <typename... Args>
    auto operator()(const Args&... args) const {
        static_assert(sizeof...(args) >= 1, "the divergence needs at least one argument");
        const auto& last = std::get<sizeof...(args) - 1>(std::forward_as_tuple(args...));
        using Last = std::remove_cv_t<std::remove_reference_t<decltype(last)>>;
        if constexpr (detail::is_face_matrix<Last>::value || detail::is_field<Last>::value) {
            using SC = typename Last::ScalarType;
            constexpr std::size_t N = Last::NUM_COMPONENTS;
            utils::Vector<NUM_TFIELDS, Data::FaceMatrixStencil<Grid, SC, N>> f;
            f[0] = FaceMatrix(last);
            for (std::size_t t = 1; t < NUM_TFIELDS; t++) f[t] = f[0];
            MultiplyAll<0>(&f, args...);
            return ApplyDivergence(f);
        } else {
            static_assert(detail::is_face_value<Last>::value,
                          "the divergence takes a FaceMatrixStencil, a Field or a FaceValueStencil last");
            Last f = last;
            MultiplyAllValues<0>(&f, args...);
            return ApplyDivergence(f);
        }
    }

    // matrix form (:141-207): neighbour +-A_d coef_nb, centre A0 cE - A0 cW + A1 cN - ... in
    // that order, times the scheme's weight of level n+1; the right-hand side is the explicit
    // part A_d (rhs_hi - rhs_lo), plus the weighted right-hand sides of the older levels
    template <typename SC, std::size_t N>
    Data::CenterMatrixStencil<Grid, SC, N> ApplyDivergence(
        const utils::Vector<NUM_TFIELDS, Data::FaceMatrixStencil<Grid, SC, N>>& s) const {
        Data::CenterMatrixStencil<Grid, SC, N> c;
        for (std::size_t n = 0; n < N; n++) {
            for (std::size_t d = 0; d < Dim; d++) {
                const auto lo = ::dare::Grid::ToCartesianNeighbor(2 * d + 1), hi = ::dare::Grid::ToCartesianNeighbor(2 * d + 2);
                c.GetValue(hi, n) = A_[d] * s[0].GetValueNeighbor(hi, n);
                if (d == 0) c.GetValue(Positions::CENTER, n) = A_[0] * s[0].GetValueCenter(hi, n);
                else c.GetValue(Positions::CENTER, n) += A_[d] * s[0].GetValueCenter(hi, n);
                c.GetValue(lo, n) = -A_[d] * s[0].GetValueNeighbor(lo, n);
                c.GetValue(Positions::CENTER, n) -= A_[d] * s[0].GetValueCenter(lo, n);
            }
        }
        c *= TimeDiscretization::template GetWeights<SC>()[0];
        for (std::size_t n = 0; n < N; n++) {
            c.GetRHS(n) = A_[0] * (s[0].GetRHS(Positions::EAST, n) - s[0].GetRHS(Positions::WEST, n));
            if constexpr (Dim > 1)
                c.GetRHS(n) += A_[1] * (s[0].GetRHS(Positions::NORTH, n) - s[0].GetRHS(Positions::SOUTH, n));
            if constexpr (Dim > 2)
                c.GetRHS(n) += A_[2] * (s[0].GetRHS(Positions::TOP, n) - s[0].GetRHS(Positions::BOTTOM, n));
        }
        for (std::size_t t = 1; t < NUM_TFIELDS; t++) {
            const SC w = TimeDiscretization::template GetWeights<SC>()[t];
            for (std::size_t n = 0; n < N; n++) {
                c.GetRHS(n) += w * A_[0] * (s[t].GetRHS(Positions::EAST, n) - s[t].GetRHS(Positions::WEST, n));
                if constexpr (Dim > 1)
                    c.GetRHS(n) += w * A_[1] * (s[t].GetRHS(Positions::NORTH, n) - s[t].GetRHS(Positions::SOUTH, n));
                if constexpr (Dim > 2)
                    c.GetRHS(n) += w * A_[2] * (s[t].GetRHS(Positions::TOP, n) - s[t].GetRHS(Positions::BOTTOM, n));
            }
        }
        return c;
    }

    // value form (:211-235): A0 E - A0 W + A1 N - A1 S + A2 T - A2 B, in that order
    template <typename SC, std::size_t N>
    utils::Vector<N, SC> ApplyDivergence(const Data::FaceValueStencil<Grid, SC, N>& s) const {
        utils::Vector<N, SC> v;
        for (std::size_t n = 0; n < N; n++) {
            v[n] = A_[0] * s.GetValue(Positions::EAST, n);
            v[n] -= A_[0] * s.GetValue(Positions::WEST, n);
            if constexpr (Dim > 1) {
                v[n] += A_[1] * s.GetValue(Positions::NORTH, n);
                v[n] -= A_[1] * s.GetValue(Positions::SOUTH, n);
            }
            if constexpr (Dim > 2) {
                v[n] += A_[2] * s.GetValue(Positions::TOP, n);
                v[n] -= A_[2] * s.GetValue(Positions::BOTTOM, n);
            }
        }
        return v;
    }

private:
    template <typename SC, std::size_t N>
    static Data::FaceMatrixStencil<Grid, SC, N> FaceMatrix(const Data::FaceMatrixStencil<Grid, SC, N>& s) {
        return s;
    }
    // a Field as the operand: unit face coefficients (:362-378)
    template <typename SC, std::size_t N>
    static Data::FaceMatrixStencil<Grid, SC, N> FaceMatrix(const Data::Field<Grid, SC, N>&) {
        Data::FaceMatrixStencil<Grid, SC, N> f;
        for (std::size_t d = 0; d < Dim; d++) {
            f.SetValues(::dare::Grid::ToCartesianNeighbor(2 * d + 1), 0, -1., 1.);
            f.SetValues(::dare::Grid::ToCartesianNeighbor(2 * d + 2), 0, 1., -1.);
        }
        for (std::size_t n = 1; n < N; n++) {
            f.GetDataCenter()[n] = f.GetDataCenter()[n - 1];
            f.GetDataNeighbor()[n] = f.GetDataNeighbor()[n - 1];
        }
        return f;
    }
    // every argument but the last multiplies each level's face matrix: a scalar scales all of
    // it, face values scale the coefficients (Divergence_Cartesian.inl:92-128, Stencils :710-717)
    template <std::size_t I, typename F, typename Arg, typename... Rest>
    static void MultiplyAll(F* f, const Arg& a, const Rest&... rest) {
        if constexpr (sizeof...(Rest) > 0) {
            for (auto& s : *f) {
                if constexpr (std::is_arithmetic_v<Arg>) s *= a;
                else s = a * s;
            }
            MultiplyAll<I + 1>(f, rest...);
        }
    }
    template <std::size_t I, typename F, typename Arg, typename... Rest>
    static void MultiplyAllValues(F* f, const Arg& a, const Rest&... rest) {
        if constexpr (sizeof...(Rest) > 0) {
            *f *= a;
            MultiplyAllValues<I + 1>(f, rest...);
        }
    }

    typename Grid::VecSC A_;
    Index ind_;
    const GridRepresentation* grep_;
};

}  // namespace dare::Matrix

namespace dare::math {

// InterpolateToFace of one component (Interpolation_Cartesian.h:59-354): the relative offset of
// the source's staggering to the target face (GetRelativeOffset, :200-219) selects 1, 2, 4 or 8
// cells (GetInterpolationIndices, :95-116), summed from the farthest to the target's own cell
// and multiplied by 1/count (InterpolateCartesianLinear, :134-151; Divisors.h:59-79).
template <std::size_t Dim, typename SC, std::size_t N>
SC InterpolateToFace(const typename Grid::Cartesian<Dim>::Representation& target,
                     const typename Grid::Cartesian<Dim>::Index& ind_target, Grid::CartesianNeighbor face,
                     const Data::GridVector<Grid::Cartesian<Dim>, SC, N>& field, std::size_t n) {
    using Index = typename Grid::Cartesian<Dim>::Index;
    auto off = field.GetGridRepresentation().GetOptions();
    const auto opt_t = target.GetOptions();
    const int fa = Grid::ToFace(face) / 2;
    off[fa] += Grid::ToNormal(face);
    Index ind = ind_target;
    Index rel;
    std::size_t aff[3] = {0, 0, 0};
    std::size_t naff = 0;
    for (std::size_t d = 0; d < Dim; d++) {
        rel[d] = off[d] - opt_t[d];
        ind[d] += rel[d] / 2;            // two half steps are a whole cell
        rel[d] = rel[d] % 2;
        if (rel[d] != 0) aff[naff++] = d;
    }
    const std::size_t count = std::size_t(1) << naff;
    Index idx[8];
    for (std::size_t p = 0; p < count; p++) idx[p] = ind;
    for (std::size_t a = 0; a < naff; a++) {
        const std::size_t step = std::size_t(1) << a, range = step * 2;
        for (std::size_t v = 0; v < count; v += range)
            for (std::size_t p = v; p < v + step; p++) idx[p][aff[a]] += rel[aff[a]];
    }
    SC value = field.At(idx[0], n);
    for (std::size_t p = 1; p < count; p++) value += field.At(idx[p], n);
    if (count > 1) value = value * (SC(1) / static_cast<SC>(count));
    return value;
}

}  // namespace dare::math

namespace dare::Matrix {

// ------------------------------------------------------------------------------ TVD
template <typename Grid, typename SC, typename FluxLimiter>
class TVD {
public:
    static constexpr std::size_t Dim = Grid::DIMENSION;
    using GridType = Grid;
    using GridRepresentation = typename Grid::Representation;
    using LO = typename Grid::LocalOrdinalType;
    using Index = typename Grid::Index;
    using Positions = typename Grid::NeighborID;

    // face velocities from one field per axis, cell-centred or staggered, through
    // InterpolateToFace (TVD_Cartesian.inl:27-38)
    TVD(const GridRepresentation& grid, LO ordinal_internal,
        utils::Vector<Dim, const Data::GridVector<Grid, SC, 1>*> v)
        : ind_(grid.MapOrdinalToIndexLocal(grid.MapInternalToLocal(ordinal_internal))) {
        for (std::size_t id = 0; id < 2 * Dim; id++) {
            const auto cnb = ::dare::Grid::ToCartesianNeighbor(id + 1);
            const SC value = math::InterpolateToFace<Dim, SC, 1>(grid, ind_, cnb, *v[id / 2], 0);
            velocity_.SetValue(cnb, 0, value);
            upwind_[id] = value >= SC(0.);
        }
    }
    // a constant velocity vector (:41-53)
    TVD(const GridRepresentation& grid, LO ordinal_internal, const utils::Vector<Dim, SC>& v)
        : ind_(grid.MapOrdinalToIndexLocal(grid.MapInternalToLocal(ordinal_internal))) {
        for (std::size_t d = 0; d < Dim; d++) {
            velocity_.SetValue(::dare::Grid::ToCartesianNeighbor(2 * d + 1), 0, v[d]);
            velocity_.SetValue(::dare::Grid::ToCartesianNeighbor(2 * d + 2), 0, v[d]);
            upwind_[2 * d] = upwind_[2 * d + 1] = v[d] >= SC(0.);
        }
    }

    // face values (:122-183): cells UU, U, D by the upwind flag, r = (U - UU) / (D - U),
    // phi_f = U + 0.5 Psi(r) (D - U)
    template <std::size_t N>
    Data::FaceValueStencil<Grid, SC, N> Interpolate(const Data::GridVector<Grid, SC, N>& field) const {
        Data::FaceValueStencil<Grid, SC, N> out;
        for (std::size_t d = 0; d < Dim; d++)
            for (int h = 0; h < 2; h++) {
                utils::Vector<N, SC> pUU, pU, pD;
                Cells(field, d, h, &pUU, &pU, &pD);
                const utils::Vector<N, SC> r = (pU - pUU) / (pD - pU);
                const utils::Vector<N, SC> psi = FluxLimiter::GetValue(r);
                const utils::Vector<N, SC> face = pU + SC(0.5) * psi * (pD - pU);
                out.SetValues(::dare::Grid::ToCartesianNeighbor(2 * d + 1 + h), face);
            }
        return out;
    }

    // implicit upwind coefficients and the deferred-correction right-hand side
    // (-0.5 Psi)(D - U) u (:185-261)
    template <std::size_t N>
    Data::FaceMatrixStencil<Grid, SC, N> operator*(const Data::GridVector<Grid, SC, N>& field) const {
        Data::FaceMatrixStencil<Grid, SC, N> s;
        utils::Vector<N, SC> ones;
        ones.SetAllValues(SC(1.));
        for (std::size_t d = 0; d < Dim; d++)
            for (int h = 0; h < 2; h++) {
                const auto face = ::dare::Grid::ToCartesianNeighbor(2 * d + 1 + h);
                const LO vp = upwind_[2 * d + h] ? 1 : 0, vn = upwind_[2 * d + h] ? 0 : 1;
                utils::Vector<N, SC> pUU, pU, pD;
                Cells(field, d, h, &pUU, &pU, &pD);
                const SC vel = velocity_.GetValue(face, 0);
                const utils::Vector<N, SC> r = (pU - pUU) / (pD - pU);
                const utils::Vector<N, SC> psi = FluxLimiter::GetValue(r);
                if (h == 0) {
                    s.SetValueNeighbor(face, SC(vp) * vel * ones);
                    s.SetValueCenter(face, SC(vn) * vel * ones);
                } else {
                    s.SetValueNeighbor(face, SC(vn) * vel * ones);
                    s.SetValueCenter(face, SC(vp) * vel * ones);
                }
                s.SetRHS(face, SC(-0.5) * psi * (pD - pU) * vel);
            }
        return s;
    }

    const Data::FaceValueStencil<Grid, SC, 1>& GetVelocities() const { return velocity_; }

private:
    // TVD_Cartesian.inl:137-143 (low face) and :162-167 (high face)
    template <std::size_t N>
    void Cells(const Data::GridVector<Grid, SC, N>& field, std::size_t d, int h, utils::Vector<N, SC>* UU,
               utils::Vector<N, SC>* U, utils::Vector<N, SC>* D) const {
        Index iUU = ind_, iU = ind_, iD = ind_;
        const LO vp = upwind_[2 * d + h] ? 1 : 0, vn = upwind_[2 * d + h] ? 0 : 1;
        if (h == 0) {
            iUU[d] += vn - 2 * vp;
            iU[d] -= vp;
            iD[d] -= vn;
        } else {
            iUU[d] += 2 * vn - vp;
            iU[d] += vn;
            iD[d] += vp;
        }
        *UU = field.GetValues(iUU);
        *U = field.GetValues(iU);
        *D = field.GetValues(iD);
    }

    Index ind_;
    Data::FaceValueStencil<Grid, SC, 1> velocity_;
    bool upwind_[2 * Dim] = {};
};

// ------------------------------------------------------------------------------ DDT
// d/dt with the scheme's weights (DDT.inl:41-81): centre = (V w0) / dt, rhs += (V w_t / dt)
// phi^{level t}; further Field arguments multiply the weights. Those property fields are read at
// the cell's internal ordinal taken as a local ordinal (DDT.inl:100), as in the reference, so
// only uniform properties give the intended result there too.
template <typename Grid, typename TimeDiscretization = timeschemes::IMPLICIT>
class DDT {
public:
    using LO = typename Grid::LocalOrdinalType;
    using SC = typename Grid::ScalarType;
    using GridRepresentation = typename Grid::Representation;
    static const std::size_t NUM_TIMESTEPS = TimeDiscretization::NUM_TIMESTEPS;
    static const std::size_t NUM_TFIELDS{NUM_TIMESTEPS + 1};

    DDT(const GridRepresentation& grep, LO ordinal, SC dt)
        : dt_(dt), ordinal_(ordinal), local_ordinal_(grep.MapInternalToLocal(ordinal)),
          volume_(grep.GetCellVolume(ordinal)) {}

    template <typename... Args>
    auto operator()(const Args&... args) const {
        static_assert(sizeof...(args) >= 1, "DDT needs the field");
        const auto& field = std::get<sizeof...(args) - 1>(std::forward_as_tuple(args...));
        using F = std::remove_cv_t<std::remove_reference_t<decltype(field)>>;
        static_assert(detail::is_field<F>::value, "the last argument of DDT is a Field");
        constexpr std::size_t N = F::NUM_COMPONENTS;
        utils::Vector<N, utils::Vector<NUM_TFIELDS, SC>> tt;
        tt[0] = volume_ * TimeDiscretization::template GetWeights<SC>() / dt_;
        for (std::size_t n = 1; n < N; n++) tt[n] = tt[n - 1];
        Properties<0>(&tt, args...);
        Data::CenterMatrixStencil<Grid, SC, N> s;
        for (std::size_t n = 0; n < N; n++) {
            s.Center(n) = tt[n][0];
            for (std::size_t t = 1; t < NUM_TFIELDS; t++) {
                const SC phi = field.GetDataVector(t).At(local_ordinal_, n);
                s.GetRHS(n) += tt[n][t] * phi;
            }
        }
        return s;
    }

private:
    template <std::size_t I, typename TT, typename Arg, typename... Rest>
    void Properties(TT* tt, const Arg& a, const Rest&... rest) const {
        if constexpr (sizeof...(Rest) > 0) {
            for (std::size_t n = 0; n < tt->size(); n++)
                for (std::size_t t = 0; t < NUM_TFIELDS; t++) (*tt)[n][t] *= a.GetDataVector(t).At(ordinal_, n);
            Properties<I + 1>(tt, rest...);
        }
    }

    SC dt_;
    LO ordinal_, local_ordinal_;
    SC volume_;
};

}  // namespace dare::Matrix
