// dare::Data stencil types of a Cartesian grid: CenterMatrixStencil, CenterValueStencil,
// FaceMatrixStencil and FaceValueStencil (Grid/Cartesian/Stencils_Cartesian.h:41-732,
// Stencils_Cartesian.inl:28-784). They carry one functor evaluation per cell on the host: the
// operator classes (dare/Operators.h) return them and MatrixBlock takes them.
//
// Positions: a center stencil holds 2 Dim + 1 entries indexed by CartesianNeighbor (CENTER = 0,
// W, E, S, N, B, T), a face stencil 2 Dim entries indexed by face (W = 0). Arithmetic is
// element-wise in the reference's order, including its two rules for stencil-by-stencil
// products: the coefficients multiply (or divide) while the right-hand sides add
// (Stencils_Cartesian.inl:133-163, 386-416), and a FaceValueStencil times a FaceMatrixStencil
// scales the centre and neighbour coefficients, not the right-hand side (:710-717).
#pragma once

#include <cstddef>

#include "Grid.h"
#include "Vector.h"

namespace dare::Data {

template <typename Grid, typename SC, std::size_t N>
class FaceValueStencil;

template <typename Grid, typename SC, std::size_t N>
class CenterMatrixStencil {
public:
    using GridType = Grid;
    using ScalarType = SC;
    static constexpr std::size_t NUM_ENTRIES = Grid::STENCIL_SIZE;
    static constexpr std::size_t NUM_COMPONENTS = N;
    using Positions = typename Grid::NeighborID;
    using ComponentArray = utils::Vector<NUM_ENTRIES, SC>;
    using DataArray = utils::Vector<N, ComponentArray>;
    using RHSType = utils::Vector<N, SC>;

    CenterMatrixStencil() = default;     // all coefficients and right-hand sides zero

    CenterMatrixStencil& operator*=(SC v) {
        for (auto& c : coefficients_) c *= v;
        for (auto& e : rhs_) e *= v;
        return *this;
    }
    CenterMatrixStencil operator*(SC v) const { return CenterMatrixStencil(*this) *= v; }
    CenterMatrixStencil& operator/=(SC v) {
        for (auto& c : coefficients_) c /= v;
        for (auto& e : rhs_) e /= v;
        return *this;
    }
    CenterMatrixStencil operator/(SC v) const { return CenterMatrixStencil(*this) /= v; }
    CenterMatrixStencil& operator+=(const CenterMatrixStencil& o) {
        coefficients_ += o.coefficients_;
        rhs_ += o.rhs_;
        return *this;
    }
    CenterMatrixStencil operator+(const CenterMatrixStencil& o) const { return CenterMatrixStencil(*this) += o; }
    CenterMatrixStencil& operator-=(const CenterMatrixStencil& o) {
        coefficients_ -= o.coefficients_;
        rhs_ -= o.rhs_;
        return *this;
    }
    CenterMatrixStencil operator-(const CenterMatrixStencil& o) const { return CenterMatrixStencil(*this) -= o; }
    CenterMatrixStencil& operator*=(const CenterMatrixStencil& o) {    // the right-hand sides add
        coefficients_ *= o.coefficients_;
        rhs_ += o.rhs_;
        return *this;
    }
    CenterMatrixStencil operator*(const CenterMatrixStencil& o) const { return CenterMatrixStencil(*this) *= o; }
    CenterMatrixStencil& operator/=(const CenterMatrixStencil& o) {
        coefficients_ /= o.coefficients_;
        rhs_ += o.rhs_;
        return *this;
    }
    CenterMatrixStencil operator/(const CenterMatrixStencil& o) const { return CenterMatrixStencil(*this) /= o; }

    SC& Center(std::size_t n) { return GetValue(Positions::CENTER, n); }
    SC Center(std::size_t n) const { return GetValue(Positions::CENTER, n); }
    void SetValue(Positions pos, std::size_t n, SC v) { coefficients_[n][::dare::Grid::ToNum(pos)] = v; }
    void SetRHS(std::size_t n, SC v) { rhs_[n] = v; }
    void SetAll(SC v) {
        for (auto& a : coefficients_) a.SetAllValues(v);
    }
    SC& GetValue(Positions pos, std::size_t n) { return coefficients_[n][::dare::Grid::ToNum(pos)]; }
    SC GetValue(Positions pos, std::size_t n) const { return coefficients_[n][::dare::Grid::ToNum(pos)]; }
    utils::Vector<N, SC> GetValues(Positions pos) const {
        utils::Vector<N, SC> v;
        for (std::size_t n = 0; n < N; n++) v[n] = GetValue(pos, n);
        return v;
    }
    DataArray& GetData() { return coefficients_; }
    const DataArray& GetData() const { return coefficients_; }
    RHSType& GetRHS() { return rhs_; }
    const RHSType& GetRHS() const { return rhs_; }
    SC& GetRHS(std::size_t n) { return rhs_[n]; }
    SC GetRHS(std::size_t n) const { return rhs_[n]; }
    CenterMatrixStencil<Grid, SC, 1> GetSlice(std::size_t n) const {
        CenterMatrixStencil<Grid, SC, 1> s;
        s.GetData()[0] = coefficients_[n];
        s.GetRHS(0) = rhs_[n];
        return s;
    }

private:
    DataArray coefficients_{};
    RHSType rhs_{};
};

template <typename Grid, typename SC, std::size_t N>
CenterMatrixStencil<Grid, SC, N> operator*(SC v, const CenterMatrixStencil<Grid, SC, N>& s) {
    return s * v;
}

// values around a cell centre (Stencils_Cartesian.h:231-271): a CenterMatrixStencil by another name
template <typename Grid, typename SC, std::size_t N>
class CenterValueStencil : public CenterMatrixStencil<Grid, SC, N> {
public:
    using MatrixStencil = CenterMatrixStencil<Grid, SC, N>;
    CenterValueStencil() = default;
    CenterValueStencil(const MatrixStencil& s) : MatrixStencil(s) {}   // NOLINT
};

template <typename Grid, typename SC, std::size_t N>
class FaceMatrixStencil;

template <typename Grid, typename SC, std::size_t N>
class FaceValueStencil {
public:
    using GridType = Grid;
    using ScalarType = SC;
    static constexpr std::size_t NUM_FACES = Grid::NUM_FACES;
    static constexpr std::size_t NUM_COMPONENTS = N;
    using Positions = typename Grid::NeighborID;
    using ComponentArray = utils::Vector<NUM_FACES, SC>;
    using DataArray = utils::Vector<N, ComponentArray>;

    FaceValueStencil() = default;

    FaceValueStencil& operator*=(SC v) {
        for (auto& c : c_) c *= v;
        return *this;
    }
    FaceValueStencil operator*(SC v) const { return FaceValueStencil(*this) *= v; }
    FaceValueStencil& operator/=(SC v) {
        for (auto& c : c_) c /= v;
        return *this;
    }
    FaceValueStencil operator/(SC v) const { return FaceValueStencil(*this) /= v; }
    FaceValueStencil& operator+=(const FaceValueStencil& o) {
        c_ += o.c_;
        return *this;
    }
    FaceValueStencil operator+(const FaceValueStencil& o) const { return FaceValueStencil(*this) += o; }
    FaceValueStencil& operator-=(const FaceValueStencil& o) {
        c_ -= o.c_;
        return *this;
    }
    FaceValueStencil operator-(const FaceValueStencil& o) const { return FaceValueStencil(*this) -= o; }
    FaceValueStencil& operator*=(const FaceValueStencil& o) {
        c_ *= o.c_;
        return *this;
    }
    FaceValueStencil operator*(const FaceValueStencil& o) const { return FaceValueStencil(*this) *= o; }
    FaceValueStencil& operator/=(const FaceValueStencil& o) {
        c_ /= o.c_;
        return *this;
    }
    FaceValueStencil operator/(const FaceValueStencil& o) const { return FaceValueStencil(*this) /= o; }
    // face values times a face matrix: centre and neighbour coefficients scale, the right-hand
    // side does not (Stencils_Cartesian.inl:710-717)
    FaceMatrixStencil<Grid, SC, N> operator*(const FaceMatrixStencil<Grid, SC, N>& m) const;

    void SetValue(Positions pos, std::size_t n, SC v) { c_[n][::dare::Grid::ToFace(pos)] = v; }
    void SetValues(Positions pos, const utils::Vector<N, SC>& v) {
        for (std::size_t n = 0; n < N; n++) c_[n][::dare::Grid::ToFace(pos)] = v[n];
    }
    void SetAll(SC v) {
        for (auto& a : c_) a.SetAllValues(v);
    }
    SC& GetValue(Positions pos, std::size_t n) { return c_[n][::dare::Grid::ToFace(pos)]; }
    SC GetValue(Positions pos, std::size_t n) const { return c_[n][::dare::Grid::ToFace(pos)]; }
    DataArray& GetData() { return c_; }
    const DataArray& GetData() const { return c_; }

private:
    DataArray c_{};
};

template <typename Grid, typename SC, std::size_t N>
FaceValueStencil<Grid, SC, N> operator*(SC v, const FaceValueStencil<Grid, SC, N>& s) {
    return s * v;
}

template <typename Grid, typename SC, std::size_t N>
class FaceMatrixStencil {
public:
    using GridType = Grid;
    using ScalarType = SC;
    static constexpr std::size_t NUM_FACES = Grid::NUM_FACES;
    static constexpr std::size_t NUM_COMPONENTS = N;
    using Positions = typename Grid::NeighborID;
    using ComponentArray = utils::Vector<NUM_FACES, SC>;
    using DataArray = utils::Vector<N, ComponentArray>;
    using RHSType = FaceValueStencil<Grid, SC, N>;

    FaceMatrixStencil() = default;

    FaceMatrixStencil& operator*=(SC v);
    FaceMatrixStencil operator*(SC v) const { return FaceMatrixStencil(*this) *= v; }
    FaceMatrixStencil& operator/=(SC v);
    FaceMatrixStencil operator/(SC v) const { return FaceMatrixStencil(*this) /= v; }
    FaceMatrixStencil& operator+=(const FaceMatrixStencil& o);
    FaceMatrixStencil operator+(const FaceMatrixStencil& o) const { return FaceMatrixStencil(*this) += o; }
    FaceMatrixStencil& operator-=(const FaceMatrixStencil& o);
    FaceMatrixStencil operator-(const FaceMatrixStencil& o) const { return FaceMatrixStencil(*this) -= o; }
    FaceMatrixStencil& operator*=(const FaceMatrixStencil& o);
    FaceMatrixStencil operator*(const FaceMatrixStencil& o) const { return FaceMatrixStencil(*this) *= o; }
    FaceMatrixStencil& operator/=(const FaceMatrixStencil& o);
    FaceMatrixStencil operator/(const FaceMatrixStencil& o) const { return FaceMatrixStencil(*this) /= o; }

    void SetValueNeighbor(Positions pos, std::size_t n, SC v) { nb_[n][::dare::Grid::ToFace(pos)] = v; }
    void SetValueCenter(Positions pos, std::size_t n, SC v) { c_[n][::dare::Grid::ToFace(pos)] = v; }
    void SetValues(Positions pos, std::size_t n, SC v_nb, SC v_c) {
        SetValueNeighbor(pos, n, v_nb);
        SetValueCenter(pos, n, v_c);
    }
    void SetRHS(Positions pos, std::size_t n, SC v);
    void SetValueNeighbor(Positions pos, const utils::Vector<N, SC>& v) {
        for (std::size_t n = 0; n < N; n++) SetValueNeighbor(pos, n, v[n]);
    }
    void SetValueCenter(Positions pos, const utils::Vector<N, SC>& v) {
        for (std::size_t n = 0; n < N; n++) SetValueCenter(pos, n, v[n]);
    }
    void SetValues(Positions pos, const utils::Vector<N, SC>& v_nb, const utils::Vector<N, SC>& v_c) {
        SetValueNeighbor(pos, v_nb);
        SetValueCenter(pos, v_c);
    }
    void SetRHS(Positions pos, const utils::Vector<N, SC>& v) {
        for (std::size_t n = 0; n < N; n++) SetRHS(pos, n, v[n]);
    }
    void SetAll(SC v) {       // the coefficients only, as in the reference (:470-478)
        for (auto& a : c_) a.SetAllValues(v);
        for (auto& a : nb_) a.SetAllValues(v);
    }
    SC& GetValueNeighbor(Positions pos, std::size_t n) { return nb_[n][::dare::Grid::ToFace(pos)]; }
    SC& GetValueCenter(Positions pos, std::size_t n) { return c_[n][::dare::Grid::ToFace(pos)]; }
    SC GetValueNeighbor(Positions pos, std::size_t n) const { return nb_[n][::dare::Grid::ToFace(pos)]; }
    SC GetValueCenter(Positions pos, std::size_t n) const { return c_[n][::dare::Grid::ToFace(pos)]; }
    SC& GetRHS(Positions pos, std::size_t n);
    SC GetRHS(Positions pos, std::size_t n) const;
    DataArray& GetDataNeighbor() { return nb_; }
    DataArray& GetDataCenter() { return c_; }
    const DataArray& GetDataNeighbor() const { return nb_; }
    const DataArray& GetDataCenter() const { return c_; }
    RHSType& GetRHS() { return rhs_; }
    const RHSType& GetRHS() const { return rhs_; }
    FaceMatrixStencil<Grid, SC, 1> GetSlice(std::size_t n) const;

private:
    DataArray nb_{}, c_{};
    RHSType rhs_{};
};

template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N> operator*(SC v, const FaceMatrixStencil<Grid, SC, N>& s) {
    return s * v;
}

// ---- out-of-line members (Stencils_Cartesian.inl:290-562, 710-717)
template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N>& FaceMatrixStencil<Grid, SC, N>::operator*=(SC v) {
    for (auto& c : nb_) c *= v;
    for (auto& c : c_) c *= v;
    rhs_ *= v;
    return *this;
}
template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N>& FaceMatrixStencil<Grid, SC, N>::operator/=(SC v) {
    for (auto& c : c_) c /= v;
    for (auto& c : nb_) c /= v;
    rhs_ /= v;
    return *this;
}
template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N>& FaceMatrixStencil<Grid, SC, N>::operator+=(const FaceMatrixStencil& o) {
    nb_ += o.nb_;
    c_ += o.c_;
    rhs_ += o.rhs_;
    return *this;
}
template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N>& FaceMatrixStencil<Grid, SC, N>::operator-=(const FaceMatrixStencil& o) {
    c_ -= o.c_;
    nb_ -= o.nb_;
    rhs_ -= o.rhs_;
    return *this;
}
template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N>& FaceMatrixStencil<Grid, SC, N>::operator*=(const FaceMatrixStencil& o) {
    c_ *= o.c_;
    nb_ *= o.nb_;
    rhs_ += o.rhs_;
    return *this;
}
template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N>& FaceMatrixStencil<Grid, SC, N>::operator/=(const FaceMatrixStencil& o) {
    c_ /= o.c_;
    nb_ /= o.nb_;
    rhs_ += o.rhs_;
    return *this;
}
template <typename Grid, typename SC, std::size_t N>
void FaceMatrixStencil<Grid, SC, N>::SetRHS(Positions pos, std::size_t n, SC v) {
    rhs_.SetValue(pos, n, v);
}
template <typename Grid, typename SC, std::size_t N>
SC& FaceMatrixStencil<Grid, SC, N>::GetRHS(Positions pos, std::size_t n) {
    return rhs_.GetValue(pos, n);
}
template <typename Grid, typename SC, std::size_t N>
SC FaceMatrixStencil<Grid, SC, N>::GetRHS(Positions pos, std::size_t n) const {
    return rhs_.GetValue(pos, n);
}
template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, 1> FaceMatrixStencil<Grid, SC, N>::GetSlice(std::size_t n) const {
    FaceMatrixStencil<Grid, SC, 1> s;
    s.GetDataNeighbor()[0] = nb_[n];
    s.GetDataCenter()[0] = c_[n];
    s.GetRHS().GetData()[0] = rhs_.GetData()[n];
    return s;
}

template <typename Grid, typename SC, std::size_t N>
FaceMatrixStencil<Grid, SC, N> FaceValueStencil<Grid, SC, N>::operator*(const FaceMatrixStencil<Grid, SC, N>& m) const {
    FaceMatrixStencil<Grid, SC, N> s(m);
    s.GetDataCenter() *= c_;
    s.GetDataNeighbor() *= c_;
    return s;
}

}  // namespace dare::Data
