// Gauge links and covariant displacement on the GPU; see gauge.hpp.
// Everything here is FP64 and per link or per (site, eigenvector): memory
// bound, small next to VdaggerV.
#include "gauge.hpp"
#include "kernels.hpp"

#include <stdexcept>
#include <string>

namespace slaph {

namespace {

struct M3 {
  double2 e[9];  // column-major: element (r, c) at e[c * 3 + r]
};

__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c_conj(double2 a) { return make_double2(a.x, -a.y); }

__device__ M3 m_load(const double2 *p) {
  M3 m;
#pragma unroll
  for (int i = 0; i < 9; ++i) m.e[i] = p[i];
  return m;
}
__device__ void m_store(double2 *p, M3 const &m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) p[i] = m.e[i];
}
__device__ M3 m_mul(M3 const &a, M3 const &b) {
  M3 c;
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = c_add(s, c_mul(a.e[k * 3 + r], b.e[col * 3 + k]));
      c.e[col * 3 + r] = s;
    }
  return c;
}
__device__ M3 m_adj(M3 const &a) {
  M3 c;
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int r = 0; r < 3; ++r) c.e[col * 3 + r] = c_conj(a.e[r * 3 + col]);
  return c;
}
__device__ M3 m_scale(M3 a, double f) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a.e[i] = make_double2(a.e[i].x * f, a.e[i].y * f);
  return a;
}
__device__ M3 m_identity() {
  M3 c;
#pragma unroll
  for (int i = 0; i < 9; ++i) c.e[i] = make_double2(i % 4 == 0 ? 1.0 : 0.0, 0.0);
  return c;
}

// exp(A) by scaling and squaring of a Taylor series: A / 2^s has 1-norm at
// most 1/4, where 18 terms leave a remainder below 2^-100 of the result.
__device__ M3 m_exp(M3 const &a) {
  double norm = 0.0;
#pragma unroll
  for (int col = 0; col < 3; ++col) {
    double cs = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) cs += hypot(a.e[col * 3 + r].x, a.e[col * 3 + r].y);
    norm = fmax(norm, cs);
  }
  int s = 0;
  while (norm > 0.25 && s < 60) {
    norm *= 0.5;
    ++s;
  }
  const M3 b = m_scale(a, ldexp(1.0, -s));
  M3 term = m_identity(), sum = m_identity();
  for (int k = 1; k <= 18; ++k) {
    term = m_scale(m_mul(term, b), 1.0 / k);
#pragma unroll
    for (int i = 0; i < 9; ++i) sum.e[i] = c_add(sum.e[i], term.e[i]);
  }
  for (int i = 0; i < s; ++i) sum = m_mul(sum, sum);
  return sum;
}

__device__ __forceinline__ double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c_scale(double2 a, double f) { return make_double2(a.x * f, a.y * f); }
__device__ M3 m_add(M3 a, M3 const &b) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a.e[i] = c_add(a.e[i], b.e[i]);
  return a;
}
__device__ double m_fro2(M3 const &a) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) s += a.e[i].x * a.e[i].x + a.e[i].y * a.e[i].y;
  return s;
}
__device__ __forceinline__ double2 m_at(M3 const &a, int r, int c) { return a.e[c * 3 + r]; }
__device__ double2 c_div(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// determinant and inverse (adjugate over determinant) of a 3 x 3 matrix
__device__ double2 m_det(M3 const &a) {
  auto m = [&](int r, int c) { return m_at(a, r, c); };
  const double2 c0 = c_sub(c_mul(m(1, 1), m(2, 2)), c_mul(m(1, 2), m(2, 1)));
  const double2 c1 = c_sub(c_mul(m(1, 2), m(2, 0)), c_mul(m(1, 0), m(2, 2)));
  const double2 c2 = c_sub(c_mul(m(1, 0), m(2, 1)), c_mul(m(1, 1), m(2, 0)));
  return c_add(c_add(c_mul(m(0, 0), c0), c_mul(m(0, 1), c1)), c_mul(m(0, 2), c2));
}
__device__ M3 m_inv(M3 const &a) {
  auto m = [&](int r, int c) { return m_at(a, r, c); };
  M3 adj;  // adj(r, c) = cofactor(c, r)
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
      adj.e[c * 3 + r] = c_sub(c_mul(m(r1, c1), m(r2, c2)), c_mul(m(r1, c2), m(r2, c1)));
    }
  const double2 d = m_det(a);
#pragma unroll
  for (int i = 0; i < 9; ++i) adj.e[i] = c_div(adj.e[i], d);
  return adj;
}
// proj_to_su3_imp (GaugeField.cpp:203-210): the unitary polar factor
// W = A (A^+ A)^(-1/2) by the Newton iteration X <- (g X + X^-+ / g) / 2 with
// g = (|X^-1|_F / |X|_F)^(1/2) (quadratic convergence from any nonsingular A;
// the loop ends when a step changes X by less than 1e-15 of its norm, at most
// 60 steps), then W det(W)^(-1/3) with the principal cube root.
__device__ M3 m_proj_su3(M3 x) {
  for (int it = 0; it < 60; ++it) {
    const M3 xi = m_inv(x);
    const double g = sqrt(sqrt(m_fro2(xi) / m_fro2(x)));
    M3 nx;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)  // (X^-+)(r, c) = conj(X^-1(c, r))
        nx.e[c * 3 + r] = c_add(c_scale(x.e[c * 3 + r], 0.5 * g), c_scale(c_conj(xi.e[r * 3 + c]), 0.5 / g));
    M3 dx = nx;
#pragma unroll
    for (int i = 0; i < 9; ++i) dx.e[i] = c_sub(nx.e[i], x.e[i]);
    x = nx;
    if (m_fro2(dx) <= 1e-30 * m_fro2(x)) break;
  }
  const double2 d = m_det(x);
  const double ang = atan2(d.y, d.x) / 3.0, mag = cbrt(hypot(d.x, d.y));
  const double2 f = make_double2(cos(ang) / mag, -sin(ang) / mag);  // 1 / (principal cube root of d)
#pragma unroll
  for (int i = 0; i < 9; ++i) x.e[i] = c_mul(x.e[i], f);
  return x;
}

__device__ __forceinline__ unsigned long long mix(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
// standard normal number k of stream (seed, i) by Box-Muller
__device__ double gauss(unsigned long long seed, unsigned long long i, int k) {
  const unsigned long long base = mix(seed ^ mix(i * 64 + unsigned(k)));
  const double u1 = ((base >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((mix(base) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ void k_su3_random(double2 *U, long long nlinks, unsigned long long seed, double eps) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nlinks) return;
  M3 x;
#pragma unroll
  for (int k = 0; k < 9; ++k) x.e[k] = make_double2(gauss(seed, i, 2 * k), gauss(seed, i, 2 * k + 1));
  // H = (X + X^+) / 2 - tr / 3: traceless Hermitian; A = i eps H
  const M3 xa = m_adj(x);
  M3 a;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 h = make_double2(0.5 * (x.e[k].x + xa.e[k].x), 0.5 * (x.e[k].y + xa.e[k].y));
    a.e[k] = make_double2(-eps * h.y, eps * h.x);
  }
  const double tr = (a.e[0].y + a.e[4].y + a.e[8].y) / 3.0;  // the trace of A is imaginary
  a.e[0].y -= tr;
  a.e[4].y -= tr;
  a.e[8].y -= tr;
  m_store(U + i * 9, m_exp(a));
}

__device__ __forceinline__ int nbr(int site, int mu, int step, int Lx, int Ly, int Lz) {
  int z = site % Lz, y = (site / Lz) % Ly, x = site / (Lz * Ly);
  if (mu == 0) x = (x + step + Lx) % Lx;
  if (mu == 1) y = (y + step + Ly) % Ly;
  if (mu == 2) z = (z + step + Lz) % Lz;
  return (x * Ly + y) * Lz + z;
}

__global__ void k_stout_step(const double2 *__restrict__ Uin, double2 *__restrict__ Uout, int Lx, int Ly, int Lz,
                             double rho) {
  const int V3 = Lx * Ly * Lz;
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= 3LL * V3) return;
  const int site = int(id / 3), mu = int(id % 3);
  auto link = [&](int s, int d) { return m_load(Uin + ((long long)s * 3 + d) * 9); };
  M3 staple;
#pragma unroll
  for (int k = 0; k < 9; ++k) staple.e[k] = make_double2(0.0, 0.0);
  for (int nu = 0; nu < 3; ++nu) {
    if (nu == mu) continue;
    // positive: U_nu(x) U_mu(x + nu) U_nu(x + mu)^+
    const int xp = nbr(site, nu, 1, Lx, Ly, Lz), xm = nbr(site, mu, 1, Lx, Ly, Lz);
    const M3 up = m_mul(link(site, nu), m_mul(link(xp, mu), m_adj(link(xm, nu))));
    // negative: U_nu(x - nu)^+ U_mu(x - nu) U_nu(x - nu + mu)
    const int xd = nbr(site, nu, -1, Lx, Ly, Lz), xdm = nbr(xd, mu, 1, Lx, Ly, Lz);
    const M3 dn = m_mul(m_adj(link(xd, nu)), m_mul(link(xd, mu), link(xdm, nu)));
#pragma unroll
    for (int k = 0; k < 9; ++k) staple.e[k] = c_add(staple.e[k], c_add(up.e[k], dn.e[k]));
  }
  const M3 u = link(site, mu);
  const M3 om = m_mul(m_scale(staple, rho / 4.0), m_adj(u));
  const M3 oa = m_adj(om);
  M3 q;
#pragma unroll
  for (int k = 0; k < 9; ++k) q.e[k] = make_double2(0.5 * (oa.e[k].x - om.e[k].x), 0.5 * (oa.e[k].y - om.e[k].y));
  const double2 tr = make_double2((oa.e[0].x - om.e[0].x + oa.e[4].x - om.e[4].x + oa.e[8].x - om.e[8].x) / 6.0,
                                  (oa.e[0].y - om.e[0].y + oa.e[4].y - om.e[4].y + oa.e[8].y - om.e[8].y) / 6.0);
  for (int d = 0; d < 3; ++d) {
    q.e[d * 4].x -= tr.x;
    q.e[d * 4].y -= tr.y;
  }
  m_store(Uout + ((long long)site * 3 + mu) * 9, m_mul(m_exp(m_scale(q, -1.0)), u));
}

// HYP step, level 1: thread per (site, mu); the two decorated links of U_mu(x)
__global__ void k_hyp_decorate(const double2 *__restrict__ Uin, double2 *__restrict__ dec, int Lx, int Ly, int Lz,
                               double alpha2) {
  const int V3 = Lx * Ly * Lz;
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= 3LL * V3) return;
  const int site = int(id / 3), mu = int(id % 3);
  auto link = [&](int s, int d) { return m_load(Uin + ((long long)s * 3 + d) * 9); };
  const M3 u = link(site, mu);
  for (int k = 0; k < 2; ++k) {
    const int rho = (mu == 0) ? 1 + k : (mu == 1 ? 2 * k : k);  // the other directions in increasing order
    const int xr = nbr(site, rho, 1, Lx, Ly, Lz), xm = nbr(site, mu, 1, Lx, Ly, Lz);
    M3 st = m_mul(link(site, rho), m_mul(link(xr, mu), m_adj(link(xm, rho))));
    const int xd = nbr(site, rho, -1, Lx, Ly, Lz), xdm = nbr(xd, mu, 1, Lx, Ly, Lz);
    st = m_add(st, m_mul(m_adj(link(xd, rho)), m_mul(link(xd, mu), link(xdm, rho))));
    const M3 a = m_add(m_scale(u, 1.0 - alpha2), m_scale(st, alpha2 / 2.0));
    m_store(dec + ((long long)site * 6 + mu + 3 * k) * 9, m_proj_su3(a));
  }
}

// slot of the link in direction d decorated in direction p (the `decor` map
// of GaugeField.cpp:167-187): d for the first other direction, d + 3 for the
// second
__device__ __forceinline__ int hyp_slot(int d, int p) {
  const int first = d == 0 ? 1 : 0;
  return p == first ? d : d + 3;
}

// HYP step, level 2: thread per (site, mu)
__global__ void k_hyp_outer(const double2 *__restrict__ Uin, const double2 *__restrict__ dec,
                            double2 *__restrict__ Uout, int Lx, int Ly, int Lz, double alpha1) {
  const int V3 = Lx * Ly * Lz;
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= 3LL * V3) return;
  const int site = int(id / 3), mu = int(id % 3);
  auto dl = [&](int s, int slot) { return m_load(dec + ((long long)s * 6 + slot) * 9); };
  M3 st;
#pragma unroll
  for (int k = 0; k < 9; ++k) st.e[k] = make_double2(0.0, 0.0);
  for (int nu = 0; nu < 3; ++nu) {
    if (nu == mu) continue;
    const int plane = 3 - mu - nu;  // the third direction
    const int a = hyp_slot(nu, plane), b = hyp_slot(mu, plane);
    const int xn = nbr(site, nu, 1, Lx, Ly, Lz), xm = nbr(site, mu, 1, Lx, Ly, Lz);
    st = m_add(st, m_mul(dl(site, a), m_mul(dl(xn, b), m_adj(dl(xm, a)))));
    const int xd = nbr(site, nu, -1, Lx, Ly, Lz), xdm = nbr(xd, mu, 1, Lx, Ly, Lz);
    st = m_add(st, m_mul(m_adj(dl(xd, a)), m_mul(dl(xd, b), dl(xdm, a))));
  }
  const M3 u = m_load(Uin + ((long long)site * 3 + mu) * 9);
  m_store(Uout + ((long long)site * 3 + mu) * 9,
          m_proj_su3(m_add(m_scale(u, 1.0 - alpha1), m_scale(st, alpha1 / 4.0))));
}

// One thread per (site, group of kDispN eigenvectors): the link matrix is
// loaded once for the group; the 64 threads of a wave take consecutive sites
// of one eigenvector, so their loads and stores of the 3-colour vectors are
// contiguous.
constexpr int kDispN = 16;

__global__ __launch_bounds__(256) void k_displace(const double2 *__restrict__ U, const double2 *__restrict__ Vin,
                                                  double2 *__restrict__ Vout, int Lx, int Ly, int Lz, int N, int mu,
                                                  int forward) {
  const int V3 = Lx * Ly * Lz;
  const int site = blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= V3) return;
  const long long K = 3LL * V3;
  const int other = nbr(site, mu, forward ? 1 : -1, Lx, Ly, Lz);
  // forward: U_mu(x) v(x + mu); backward: U_mu(x - mu)^+ v(x - mu)
  const double2 *u = U + ((long long)(forward ? site : other) * 3 + mu) * 9;
  double2 m[9];  // the applied 3 x 3 matrix, column-major
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c * 3 + r] = forward ? u[c * 3 + r] : c_conj(u[r * 3 + c]);
  const int n0 = blockIdx.y * kDispN, n1 = min(N, n0 + kDispN);
  for (int n = n0; n < n1; ++n) {
    const double2 *v = Vin + n * K + 3LL * other;
    const double2 v0 = v[0], v1 = v[1], v2 = v[2];
    double2 *o = Vout + n * K + 3LL * site;
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] = c_add(c_add(c_mul(m[r], v0), c_mul(m[3 + r], v1)), c_mul(m[6 + r], v2));
  }
}

// Spatial plaquette sum: thread per site, the six ordered planes (mu, nu),
// Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+) / 3 summed per site, then
// per block (fixed tree order) into partial[blockIdx.x].
__global__ __launch_bounds__(256) void k_plaquette(const double2 *__restrict__ U, int Lx, int Ly, int Lz,
                                                   double *__restrict__ partial) {
  __shared__ double red[256];
  const int V3 = Lx * Ly * Lz;
  const int site = blockIdx.x * 256 + threadIdx.x;
  double p = 0.0;
  if (site < V3) {
    auto link = [&](int s, int d) { return m_load(U + ((long long)s * 3 + d) * 9); };
    for (int mu = 0; mu < 3; ++mu)
      for (int nu = 0; nu < 3; ++nu) {
        if (mu == nu) continue;
        const int j = nbr(site, mu, 1, Lx, Ly, Lz), l = nbr(site, nu, 1, Lx, Ly, Lz);
        const M3 a = m_mul(link(site, mu), link(j, nu));
        const M3 b = m_mul(m_adj(link(l, mu)), m_adj(link(site, nu)));
        // Re tr(a b) = sum_{r,c} Re(a[r][c] b[c][r])
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            const double2 x = m_at(a, r, c), y = m_at(b, c, r);
            p += x.x * y.x - x.y * y.y;
          }
      }
    p /= 3.0;
  }
  red[threadIdx.x] = p;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

dim3 blocks_for(long long n) { return dim3(unsigned((n + 255) / 256)); }

}  // namespace

int plaquette_blocks(int Lx, int Ly, int Lz) { return int(((long long)Lx * Ly * Lz + 255) / 256); }

void launch_plaquette(const double2 *U, int Lx, int Ly, int Lz, double *partial, hipStream_t s) {
  const long long V3 = (long long)Lx * Ly * Lz;
  if (V3 <= 0) return;
  if ((V3 + 255) / 256 > 0x7fffffffLL) throw std::runtime_error("launch_plaquette: grid too large");
  hipLaunchKernelGGL(k_plaquette, blocks_for(V3), dim3(256), 0, s, U, Lx, Ly, Lz, partial);
  SLAPH_HIP(hipGetLastError());
}

void launch_su3_random(double2 *U, long long nlinks, unsigned long long seed, double eps, hipStream_t s) {
  if (nlinks <= 0) return;
  if ((nlinks + 255) / 256 > 0x7fffffffLL) throw std::runtime_error("launch_su3_random: too many links");
  hipLaunchKernelGGL(k_su3_random, blocks_for(nlinks), dim3(256), 0, s, U, nlinks, seed, eps);
  SLAPH_HIP(hipGetLastError());
}

void launch_stout_step(const double2 *Uin, double2 *Uout, int Lx, int Ly, int Lz, double rho, hipStream_t s) {
  const long long n = 3LL * Lx * Ly * Lz;
  if (n <= 0) return;
  if (Uin == Uout) throw std::runtime_error("launch_stout_step: the step needs separate input and output links");
  hipLaunchKernelGGL(k_stout_step, blocks_for(n), dim3(256), 0, s, Uin, Uout, Lx, Ly, Lz, rho);
  SLAPH_HIP(hipGetLastError());
}

void launch_hyp_decorate(const double2 *U, double2 *dec, int Lx, int Ly, int Lz, double alpha2, hipStream_t s) {
  const long long n = 3LL * Lx * Ly * Lz;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hyp_decorate, blocks_for(n), dim3(256), 0, s, U, dec, Lx, Ly, Lz, alpha2);
  SLAPH_HIP(hipGetLastError());
}

void launch_hyp_outer(const double2 *U, const double2 *dec, double2 *Uout, int Lx, int Ly, int Lz, double alpha1,
                      hipStream_t s) {
  const long long n = 3LL * Lx * Ly * Lz;
  if (n <= 0) return;
  if (U == Uout) throw std::runtime_error("launch_hyp_outer: the step needs separate input and output links");
  hipLaunchKernelGGL(k_hyp_outer, blocks_for(n), dim3(256), 0, s, U, dec, Uout, Lx, Ly, Lz, alpha1);
  SLAPH_HIP(hipGetLastError());
}

void launch_displace(const double2 *U, const double2 *Vin, double2 *Vout, int Lx, int Ly, int Lz, int N, int mu,
                     int forward, hipStream_t s) {
  const long long n = (long long)Lx * Ly * Lz * N;
  if (n <= 0) return;
  if (mu < 0 || mu > 2) throw std::runtime_error("launch_displace: direction must be 0, 1 or 2");
  if (Vin == Vout) throw std::runtime_error("launch_displace: the displacement needs separate input and output");
  const long long V3 = (long long)Lx * Ly * Lz;
  if ((V3 + 255) / 256 > 0x7fffffffLL || (N + kDispN - 1) / kDispN > 65535)
    throw std::runtime_error("launch_displace: grid too large");
  hipLaunchKernelGGL(k_displace, dim3(unsigned((V3 + 255) / 256), unsigned((N + kDispN - 1) / kDispN)), dim3(256), 0, s,
                     U, Vin, Vout, Lx, Ly, Lz, N, mu, forward);
  SLAPH_HIP(hipGetLastError());
}

}  // namespace slaph
