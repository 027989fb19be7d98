// Gauge links of one timeslice and the covariant displacement of eigenvectors
// (reference: src/GaugeField.cpp, src/OperatorFactory.cpp:33-55).
//
// Layout: U[(site * 3 + mu) * 9 + c * 3 + r] = element (r, c) of the 3 x 3
// link U_mu(site) (column-major, the element order of Eigen::Matrix3cd),
// site = (x Ly + y) Lz + z, mu = 0, 1, 2 for x, y, z.  Eigenvector matrices
// are K x N column-major with row 3 site + colour.
#pragma once

#include <hip/hip_runtime.h>

namespace slaph {

// Seeded SU(3) links U = exp(i eps H), H a traceless Hermitian matrix from
// complex Gaussian entries (stream `seed`, link index as counter).
void launch_su3_random(double2 *U, long long nlinks, unsigned long long seed, double eps, hipStream_t s);
// One stout step of smearing_stout (GaugeField.cpp:228-274) on a timeslice:
// staple C = sum_{nu != mu} U_nu(x) U_mu(x+nu) U_nu(x+mu)^+ +
// U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu), Omega = (rho / 4) C U_mu(x)^+,
// q = (Omega^+ - Omega) / 2 - tr(Omega^+ - Omega) / 6, U' = exp(-q) U.
void launch_stout_step(const double2 *Uin, double2 *Uout, int Lx, int Ly, int Lz, double rho, hipStream_t s);
// One HYP step of smearing_hyp (GaugeField.cpp:322-431) on a timeslice, in
// two launches.  launch_hyp_decorate: for every link U_mu(x) and each of the
// two other directions rho, the decorated link
// P[(1 - alpha2) U_mu(x) + alpha2 / 2 (U_rho(x) U_mu(x+rho) U_rho(x+mu)^+ +
// U_rho(x-rho)^+ U_mu(x-rho) U_rho(x-rho+mu))] into dec[(x * 6 + slot) * 9],
// slot mu for the first other direction and mu + 3 for the second.
// launch_hyp_outer: U'_mu(x) = P[(1 - alpha1) U_mu(x) + alpha1 / 4 (sum over
// nu != mu of the staples of the links decorated in the third direction)].
// P is proj_to_su3_imp: A (A^+ A)^(-1/2) (the unitary polar factor, by the
// scaled Newton iteration) times det^(-1/3) (principal root).
void launch_hyp_decorate(const double2 *U, double2 *dec, int Lx, int Ly, int Lz, double alpha2, hipStream_t s);
void launch_hyp_outer(const double2 *U, const double2 *dec, double2 *Uout, int Lx, int Ly, int Lz, double alpha1,
                      hipStream_t s);
// One one-link displacement of the N columns of V (K = 3 Lx Ly Lz rows):
// forward (the reference's '>'): out(x) = U_mu(x) V(x + mu);
// backward ('<'): out(x) = U_mu(x - mu)^+ V(x - mu).
void launch_displace(const double2 *U, const double2 *Vin, double2 *Vout, int Lx, int Ly, int Lz, int N, int mu,
                     int forward, hipStream_t s);
// Spatial plaquette of a timeslice (GaugeField::plaque_ts, GaugeField.cpp:
// 722-760): per 256-site block the sum over its sites and the six ordered
// planes mu != nu of Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+) / 3
// into partial[block], plaquette_blocks(...) entries; the plaquette is their
// sum / (6 Lx Ly Lz).
int plaquette_blocks(int Lx, int Ly, int Lz);
void launch_plaquette(const double2 *U, int Lx, int Ly, int Lz, double *partial, hipStream_t s);

}  // namespace slaph
