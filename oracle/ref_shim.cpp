/*
 * Test infrastructure: exposes the reference's own standalone headers (compiled from
 * /root/reference by oracle/Makefile into oracle/_ref/) as C functions, so the oracle's
 * restatement can be checked against the reference code itself where that code builds
 * without Trilinos/Kokkos:
 *   Equations/FluxLimiter.h          (CDS, UPWIND, VANALBADA, MINMOD)
 *   AnalyticalSolutions/Diffusion.h  (Diffusion_N0D1)
 *   Utilities/Vector.h               (element-wise operators used by every stencil)
 */
#include <cstddef>
#include <type_traits>

#include "Utilities/Vector.h"
#include "Equations/FluxLimiter.h"
#include "AnalyticalSolutions/Diffusion.h"

namespace {
template <typename L>
dare::utils::Vector<1, double> lim(const dare::utils::Vector<1, double>& r) {
    return L::GetValue(r);
}
dare::utils::Vector<1, double> lim_id(int id, const dare::utils::Vector<1, double>& r) {
    switch (id) {
    case 0: return lim<dare::Matrix::CDS>(r);
    case 1: return lim<dare::Matrix::UPWIND>(r);
    case 2: return lim<dare::Matrix::VANALBADA>(r);
    default: return lim<dare::Matrix::MINMOD>(r);
    }
}
}  // namespace

extern "C" {
double ref_limiter(int id, double r) {
    dare::utils::Vector<1, double> v;
    v[0] = r;
    return lim_id(id, v)[0];
}

double ref_diffusion_n0d1(double tau, double zeta, int n_modes) {
    return dare::analytical::Diffusion_N0D1(tau, zeta, n_modes);
}

/* The TVD face value and deferred-correction expressions of TVD_Cartesian.inl:152-157 and
 * :215-219, evaluated with the reference's Vector operators and limiter types. */
double ref_tvd_face_value(int id, double uu, double u, double d) {
    dare::utils::Vector<1, double> phi_UU, phi_U, phi_D;
    phi_UU[0] = uu;
    phi_U[0] = u;
    phi_D[0] = d;
    auto r_f = (phi_U - phi_UU) / (phi_D - phi_U);
    auto flux_lim = lim_id(id, r_f);
    auto phi_face = phi_U + 0.5 * flux_lim * (phi_D - phi_U);
    return phi_face[0];
}

double ref_tvd_face_rhs(int id, double uu, double u, double d, double vel) {
    dare::utils::Vector<1, double> phi_UU, phi_U, phi_D;
    phi_UU[0] = uu;
    phi_U[0] = u;
    phi_D[0] = d;
    auto r_f = (phi_U - phi_UU) / (phi_D - phi_U);
    auto flux_lim = lim_id(id, r_f);
    auto phi_explicit = -0.5 * flux_lim * (phi_D - phi_U) * vel;
    return phi_explicit[0];
}
}
