// sphdem_adjoint.hip -- forward and backward (vector-Jacobian product)
// kernels of the pair passes of the differentiable SPH-DEM step: SPH force,
// omega, dddt, the fluid-side coupling sums (porosity, Shepard sums,
// DEM->SPH drag), fluid-particle interaction with Di Felice drag, DEM contacts.
//
// The forward kernels restate the reference operators over explicit pair sets
// in original particle order (sphIncompress.h:899-972 with sph.h:579-666,
// 692-734; sph.h:1708-1792 with 1575-1631; sph.h:1676-1706 with 1495-1516).
// Each backward kernel recomputes its pair's forward terms and propagates the
// output adjoints to every differentiable input by the chain rule written out
// by hand (derivations in the comments).  Sums over pairs go through hardware
// FP64 atomics on global memory; the parameter adjoints are reduced per block
// first, so one atomic per block lands on each.
//
// Wendland C2 (3D, support 2h) in the closed forms used below, with q = r/h,
// t = 2 - q and c = WCON:
//   K(q) = c t^4 (1 + 2q),   K'(q) = -10 c q t^3
//   W    = K / h^3           dW/dr = K'/h^4      dW/dh = -(q K' + 3 K)/h^4
//   F    = K'/(q h^5) = -10 c t^3 / h^5  (the gradW factor, sph.h:74-81)
//   dF/dr = 30 c t^2 / h^6   dF/dh = -10 c t^2 (3q - 5t) / h^6

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "../../include/sphdem_adjoint.h"
#include "../../include/sphdem_c.h"

namespace sphdem_adj {

constexpr int kTPB = 256;
constexpr int SPH_T = 0;

// Thread index with the grid's blocks dealt to the 8 XCDs in contiguous
// ranges (the hardware deals consecutive blocks round-robin): consecutive
// queries are neighbours in space, so each XCD's L2 then holds the partner
// data of its own range instead of every XCD fetching all of it.  A bijection
// on the grid.
__device__ __forceinline__ int xcdIndex() {
  const int nb = gridDim.x, x = blockIdx.x & 7, r = blockIdx.x >> 3;
  const int q = nb >> 3, rem = nb & 7;
  return (x * q + min(x, rem) + r) * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ double3 ld3(const double* p, int i) {
  return make_double3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ __forceinline__ void add3(double* p, int i, double x, double y, double z) {
  unsafeAtomicAdd(p + 3 * i, x);
  unsafeAtomicAdd(p + 3 * i + 1, y);
  unsafeAtomicAdd(p + 3 * i + 2, z);
}
__device__ __forceinline__ void add1(double* p, int i, double x) { unsafeAtomicAdd(p + i, x); }

// b's periodic image: r_b + shift
__device__ __forceinline__ double3 imageOf(const double* r, int b, int img, const double* L) {
  const double3 rb = ld3(r, b);
  return make_double3(rb.x + (double)(img % 3 - 1) * L[0], rb.y + (double)((img / 3) % 3 - 1) * L[1],
                      rb.z + (double)(img / 9 - 1) * L[2]);
}

__device__ __forceinline__ void wendF(double c, double r, double h, double& F, double& dFdr,
                                      double& dFdh) {
  const double q = r / h;
  if (q <= 2.0) {
    const double t = 2.0 - q, ih = 1.0 / h, ih2 = ih * ih, ih5 = ih2 * ih2 * ih, ih6 = ih5 * ih;
    F = -10.0 * c * t * t * t * ih5;
    dFdr = 30.0 * c * t * t * ih6;
    dFdh = -10.0 * c * t * t * (3.0 * q - 5.0 * t) * ih6;
  } else {
    F = dFdr = dFdh = 0.0;
  }
}
__device__ __forceinline__ void wendW(double c, double r, double h, double& W, double& dWdr,
                                      double& dWdh) {
  const double q = r / h;
  if (q <= 2.0) {
    const double t = 2.0 - q, ih = 1.0 / h, ih3 = ih * ih * ih, ih4 = ih3 * ih;
    const double K = c * t * t * t * t * (1.0 + 2.0 * q);
    const double Kp = -10.0 * c * q * t * t * t;
    W = K * ih3;
    dWdr = Kp * ih4;
    dWdh = -(q * Kp + 3.0 * K) * ih4;
  } else {
    W = dWdr = dWdh = 0.0;
  }
}

// Sums NV per-thread values over the block and adds each total to dst[k]
// with one atomic.  Every thread of the block must call it.
template <int NV>
__device__ void blockAdd(double (&v)[NV], double* dst) {
  __shared__ double part[NV][kTPB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) part[k][w] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int j = 0; j < kTPB / 64; j++) s += part[threadIdx.x][j];
    if (s != 0.0) unsafeAtomicAdd(dst + threadIdx.x, s);
  }
}

// ---------------------------------------------------------------- SPH force
// Query a over its pairs (a, b), dx = r_a - r_b', r = |dx|, Fa = F(r, h_a),
// Fb = F(r, h_b), ia = 1/omega_a:
//   viscous (a SPH):  A  = m_b * visc * d / den,  fv_a += A (vhat_a - vhat_b)
//     visc = 2 nu rho0 / (rho_a rho_b),  d = r^2 (Fa + Fb) ia / 2,
//     den = r^2 + 0.01 hav^2,  hav = (h_a + h_b)/2
//   pressure:         B  = -m_b (pdr2_a Fa + pdr2_b Fb) ia,  fp_a += B dx
// One thread per query particle: its sums stay in registers.
__global__ void __launch_bounds__(kTPB) k_force_fwd(int32_t nq, const int32_t* __restrict__ query,
                            const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                            const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                            sphdem_adj_consts C, double* __restrict__ fv,
                            double* __restrict__ fp) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double3 ra = ld3(in.r, a), va = ld3(in.vhat, a);
  const double ha = in.h[a], ia = 1.0 / in.omega[a], rhoa = in.dens[a], pda = in.pdr2[a];
  const bool sphA = in.iam[a] == SPH_T;
  double vx = 0.0, vy = 0.0, vz = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int b = pb[j];
    const double3 rb = imageOf(in.r, b, pimg[j], C.shift_len);
    const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
    const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
    const double hb = in.h[b];
    double Fa, Fb, t0, t1;
    wendF(C.wcon, r, ha, Fa, t0, t1);
    wendF(C.wcon, r, hb, Fb, t0, t1);
    const double mb = in.mass[b];
    if (sphA) {
      const double visc = 2.0 * C.viscosity * C.dens / (rhoa * in.dens[b]);
      const double d = r2 * (Fa + Fb) * 0.5 * ia;
      const double hav = 0.5 * (ha + hb);
      const double den = r2 + 0.01 * hav * hav;
      const double A = mb * visc * d / den;
      const double3 vb = ld3(in.vhat, b);
      vx += A * (va.x - vb.x);
      vy += A * (va.y - vb.y);
      vz += A * (va.z - vb.z);
    }
    const double B = -mb * (pda * Fa + in.pdr2[b] * Fb) * ia;
    px += B * dx;
    py += B * dy;
    pz += B * dz;
  }
  fv[3 * a] += vx;
  fv[3 * a + 1] += vy;
  fv[3 * a + 2] += vz;
  fp[3 * a] += px;
  fp[3 * a + 1] += py;
  fp[3 * a + 2] += pz;
}

// Backward over the symmetric SPH lists, gathered: the thread of a adds the
// adjoints of a's inputs from its own force terms (a the query) and from the
// terms of its partners' forces in which a is the partner (b the query: the
// same r, dx and dv with their signs flipped, F taken on both h, divided by
// omega_b), so every adjoint is written by one thread, without atomics.  The
// viscosity adjoint is summed over the own terms only (each pair once).
__global__ void __launch_bounds__(kTPB) k_force_bwd(int32_t nq, const int32_t* __restrict__ query,
                            const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                            const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                            sphdem_adj_consts C, const double* __restrict__ gfv,
                            const double* __restrict__ gfp, sphdem_adj_grads g) {
  const int i = xcdIndex();
  double gnu[1] = {0.0};
  if (i < nq) {
    const int a = query[i];
    const double3 ra = ld3(in.r, a), va = ld3(in.vhat, a);
    const double ha = in.h[a], ia = 1.0 / in.omega[a], rhoa = in.dens[a], pda = in.pdr2[a];
    const double ma = in.mass[a];
    const bool sphA = in.iam[a] == SPH_T;
    const double3 gv = ld3(gfv, a), gp = ld3(gfp, a);
    double grx = 0.0, gry = 0.0, grz = 0.0, gvx = 0.0, gvy = 0.0, gvz = 0.0;
    double gha = 0.0, gda = 0.0, gia = 0.0, gpa = 0.0;
    for (int64_t j = off[i]; j < off[i + 1]; j++) {
      const int b = pb[j];
      const double3 rb = imageOf(in.r, b, pimg[j], C.shift_len);
      const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
      const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
      const double hb = in.h[b], rhob = in.dens[b], pdb = in.pdr2[b], mb = in.mass[b];
      double Fa, dFa_r, dFa_h, Fb, dFb_r, dFb_h;
      wendF(C.wcon, r, ha, Fa, dFa_r, dFa_h);
      wendF(C.wcon, r, hb, Fb, dFb_r, dFb_h);
      const double3 vb = ld3(in.vhat, b);
      const double dvx = va.x - vb.x, dvy = va.y - vb.y, dvz = va.z - vb.z;
      const double visc = 2.0 * C.viscosity * C.dens / (rhoa * rhob);
      const double Fs = Fa + Fb;
      const double hav = 0.5 * (ha + hb);
      const double den = r2 + 0.01 * hav * hav;
      // adjoints gathered for this pair: of r^2, F(r, h_a), F(r, h_b), and
      // the direct dx term; a's own h, dens, pdr2 in the running sums
      double r2bar = 0.0, Fabar = 0.0, Fbbar = 0.0;
      double dgx = 0.0, dgy = 0.0, dgz = 0.0;
      // ---- a's own terms (a the query)
      if (sphA) {
        const double d = r2 * Fs * 0.5 * ia;
        const double A = mb * visc * d / den;
        const double Abar = gv.x * dvx + gv.y * dvy + gv.z * dvz;  // fv_a = A dv
        gvx += A * gv.x;
        gvy += A * gv.y;
        gvz += A * gv.z;
        const double viscbar = Abar * mb * d / den;
        const double dbar = Abar * mb * visc / den;
        const double denbar = -Abar * A / den;
        gnu[0] += viscbar * visc / C.viscosity;
        gda += -viscbar * visc / rhoa;
        r2bar += dbar * Fs * 0.5 * ia + denbar;
        const double Fsbar = dbar * r2 * 0.5 * ia;
        gia += dbar * r2 * Fs * 0.5;
        Fabar += Fsbar;
        Fbbar += Fsbar;
        gha += 0.5 * (denbar * 0.02 * hav);
      }
      {  // pressure: fp_a = B dx, B = -m_b (pdr2_a Fa + pdr2_b Fb) ia
        const double B = -mb * (pda * Fa + pdb * Fb) * ia;
        const double Bbar = gp.x * dx + gp.y * dy + gp.z * dz;
        gpa += -mb * Fa * ia * Bbar;
        Fabar += -mb * pda * ia * Bbar;
        Fbbar += -mb * pdb * ia * Bbar;
        gia += -mb * (pda * Fa + pdb * Fb) * Bbar;
        dgx += B * gp.x;
        dgy += B * gp.y;
        dgz += B * gp.z;
      }
      // ---- b's terms with a as the partner (dx, dv flip sign; ib = 1/omega_b)
      const double3 gvb = ld3(gfv, b), gpb = ld3(gfp, b);
      const double ib = 1.0 / in.omega[b];
      if (in.iam[b] == SPH_T) {
        const double d = r2 * Fs * 0.5 * ib;
        const double A = ma * visc * d / den;
        // fv_b = A (vhat_b - vhat_a) = -A dv
        const double Abar = -(gvb.x * dvx + gvb.y * dvy + gvb.z * dvz);
        gvx += -A * gvb.x;
        gvy += -A * gvb.y;
        gvz += -A * gvb.z;
        const double viscbar = Abar * ma * d / den;
        const double dbar = Abar * ma * visc / den;
        const double denbar = -Abar * A / den;
        gda += -viscbar * visc / rhoa;
        r2bar += dbar * Fs * 0.5 * ib + denbar;
        const double Fsbar = dbar * r2 * 0.5 * ib;
        Fabar += Fsbar;
        Fbbar += Fsbar;
        gha += 0.5 * (denbar * 0.02 * hav);
      }
      {  // fp_b = B (-dx), B = -m_a (pdr2_b Fb + pdr2_a Fa) ib
        const double B = -ma * (pdb * Fb + pda * Fa) * ib;
        const double Bbar = -(gpb.x * dx + gpb.y * dy + gpb.z * dz);
        gpa += -ma * Fa * ib * Bbar;
        Fabar += -ma * pda * ib * Bbar;
        Fbbar += -ma * pdb * ib * Bbar;
        // d fp_b / d r_a through -dx: -B gp_b
        dgx += -B * gpb.x;
        dgy += -B * gpb.y;
        dgz += -B * gpb.z;
      }
      // F(r, h): the r dependence goes to r_a; the h_a dependence of F(r, h_a)
      const double rbar = Fabar * dFa_r + Fbbar * dFb_r;
      gha += Fabar * dFa_h;
      const double sc = 2.0 * r2bar + (r > 0.0 ? rbar / r : 0.0);
      grx += dgx + sc * dx;
      gry += dgy + sc * dy;
      grz += dgz + sc * dz;
    }
    g.r[3 * a] += grx;
    g.r[3 * a + 1] += gry;
    g.r[3 * a + 2] += grz;
    g.vhat[3 * a] += gvx;
    g.vhat[3 * a + 1] += gvy;
    g.vhat[3 * a + 2] += gvz;
    g.h[a] += gha;
    g.dens[a] += gda;
    g.omega[a] += -gia * ia * ia;  // ia = 1/omega_a
    g.pdr2[a] += gpa;
  }
  blockAdd<1>(gnu, g.params);
}

// ---------------------------------------------------------------- omega
// Nsph::calcOmega (sph.h:437-452): omega_a = 1 + h_a S_a / (3 rho_a),
// S_a = m_a P(0, h_a) + sum_b m_b P(r, h_a), P = dW/dh = G(q)/h^4 with
//   G = 10 c q^2 t^3 - 3 K,  G' = 50 c q t^3 - 30 c q^2 t^2,
//   dP/dr = G'/h^5,  dP/dh = -(q G' + 4 G)/h^5.
__device__ __forceinline__ void wendP(double c, double r, double h, double& P, double& dPdr,
                                      double& dPdh) {
  const double q = r / h;
  if (q <= 2.0) {
    const double t = 2.0 - q, ih = 1.0 / h, ih4 = (ih * ih) * (ih * ih), ih5 = ih4 * ih;
    const double t2 = t * t, t3 = t2 * t;
    const double G = 10.0 * c * q * q * t3 - 3.0 * c * t3 * t * (1.0 + 2.0 * q);
    const double Gp = 50.0 * c * q * t3 - 30.0 * c * q * q * t2;
    P = G * ih4;
    dPdr = Gp * ih5;
    dPdh = -(q * Gp + 4.0 * G) * ih5;
  } else {
    P = dPdr = dPdh = 0.0;
  }
}

__global__ void __launch_bounds__(kTPB) k_omega_fwd(int32_t nq, const int32_t* __restrict__ query,
                            const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                            const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                            sphdem_adj_consts C, double* __restrict__ omega) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double3 ra = ld3(in.r, a);
  const double ha = in.h[a];
  double P, t0, t1;
  wendP(C.wcon, 0.0, ha, P, t0, t1);
  double S = in.mass[a] * P;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int b = pb[j];
    const double3 rb = imageOf(in.r, b, pimg[j], C.shift_len);
    const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
    wendP(C.wcon, sqrt(dx * dx + dy * dy + dz * dz), ha, P, t0, t1);
    S += in.mass[b] * P;
  }
  omega[a] += ha * S / (in.dens[a] * 3.0);
}

// Backward of omega over the symmetric SPH lists (b in the list of a iff a is
// in the list of b, with the opposite image): the thread of a adds the terms
// of its own sum and, from the same entries, those of its partners' sums in
// which a appears, so every adjoint is written by one thread, without atomics.
//   own:     Sbar_a = go_a h_a / (3 rho_a),  r_a += Sbar_a m_b dP(r, h_a)/dr dx/r
//   partner: r_a   += Sbar_b m_a dP(r, h_b)/dr dx/r     (dx = r_a - r_b')
__global__ void __launch_bounds__(kTPB) k_omega_bwd(int32_t nq, const int32_t* __restrict__ query,
                            const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                            const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                            sphdem_adj_consts C, const double* __restrict__ gomega,
                            sphdem_adj_grads g) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double go = gomega[a];
  const double3 ra = ld3(in.r, a);
  const double ha = in.h[a], rhoa = in.dens[a], ma = in.mass[a];
  double P0, dP0r, dP0h, P, dPr, dPh;
  wendP(C.wcon, 0.0, ha, P0, dP0r, dP0h);
  double S = ma * P0;
  if (go != 0.0) {
    for (int64_t j = off[i]; j < off[i + 1]; j++) {
      const int b = pb[j];
      const double3 rb = imageOf(in.r, b, pimg[j], C.shift_len);
      const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
      wendP(C.wcon, sqrt(dx * dx + dy * dy + dz * dz), ha, P, dPr, dPh);
      S += in.mass[b] * P;
    }
  }
  const double Sbar = go * ha / (3.0 * rhoa);
  double gha = go * S / (3.0 * rhoa) + Sbar * ma * dP0h;
  const double gda = -go * ha * S / (3.0 * rhoa * rhoa);
  double grx = 0.0, gry = 0.0, grz = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int b = pb[j];
    const double3 rb = imageOf(in.r, b, pimg[j], C.shift_len);
    const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double k = 0.0;
    if (Sbar != 0.0) {  // a's own sum
      wendP(C.wcon, r, ha, P, dPr, dPh);
      const double c = Sbar * in.mass[b];
      gha += c * dPh;
      k += c * dPr;
    }
    const double gob = gomega[b];
    if (gob != 0.0) {  // b's sum, in which a is the partner
      const double hb = in.h[b];
      double Pb, dPbr, dPbh;
      wendP(C.wcon, r, hb, Pb, dPbr, dPbh);
      k += gob * hb / (3.0 * in.dens[b]) * ma * dPbr;
    }
    if (r > 0.0 && k != 0.0) {
      const double kr = k / r;
      grx += kr * dx;
      gry += kr * dy;
      grz += kr * dz;
    }
  }
  g.r[3 * a] += grx;
  g.r[3 * a + 1] += gry;
  g.r[3 * a + 2] += grz;
  g.h[a] += gha;
  g.dens[a] += gda;
}

// ---------------------------------------------------------------- dddt
// CsphIncompress::calcDddtDudt (sphIncompress.h:783-851) with gradW on h_a
// and omega_a: dddt_a = sum_b m_b (vhat_a - vhat_b).dx F(r, h_a) / omega_a.
__global__ void __launch_bounds__(kTPB) k_dddt_fwd(int32_t nq, const int32_t* __restrict__ query,
                           const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                           const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                           sphdem_adj_consts C, double* __restrict__ dddt) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double3 ra = ld3(in.r, a), va = ld3(in.vhat, a);
  const double ha = in.h[a], ia = 1.0 / in.omega[a];
  double sum = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int b = pb[j];
    const double3 rb = imageOf(in.r, b, pimg[j], C.shift_len);
    const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double F, t0, t1;
    wendF(C.wcon, r, ha, F, t0, t1);
    const double3 vb = ld3(in.vhat, b);
    const double u = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
    sum += in.mass[b] * u * F * ia;
  }
  dddt[a] += sum;
}

// Backward of dddt over the symmetric SPH lists, gathered as for omega: the
// thread of a adds the terms of dddt_a and of every dddt_b in which a is the
// partner (u = dv.dx and dx change sign between the two, F is taken on the
// query's h, divided by the query's omega).
__global__ void __launch_bounds__(kTPB) k_dddt_bwd(int32_t nq, const int32_t* __restrict__ query,
                           const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                           const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                           sphdem_adj_consts C, const double* __restrict__ gdddt,
                           sphdem_adj_grads g) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double gd = gdddt[a];
  const double3 ra = ld3(in.r, a), va = ld3(in.vhat, a);
  const double ha = in.h[a], ia = 1.0 / in.omega[a], ma = in.mass[a];
  double grx = 0.0, gry = 0.0, grz = 0.0, gvx = 0.0, gvy = 0.0, gvz = 0.0, gha = 0.0, gia = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int b = pb[j];
    const double gdb = gdddt[b];
    if (gd == 0.0 && gdb == 0.0) continue;
    const double3 rb = imageOf(in.r, b, pimg[j], C.shift_len);
    const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double3 vb = ld3(in.vhat, b);
    const double dvx = va.x - vb.x, dvy = va.y - vb.y, dvz = va.z - vb.z;
    const double u = dvx * dx + dvy * dy + dvz * dz;  // the same for (b, a)
    double ubar = 0.0, k = 0.0;
    if (gd != 0.0) {  // term m_b u F(r, h_a) / omega_a of dddt_a
      double F, dFr, dFh;
      wendF(C.wcon, r, ha, F, dFr, dFh);
      const double mb = in.mass[b];
      ubar += gd * mb * F * ia;
      const double Fbar = gd * mb * u * ia;
      gia += gd * mb * u * F;
      gha += Fbar * dFh;
      k += Fbar * dFr;
    }
    if (gdb != 0.0) {  // term m_a u F(r, h_b) / omega_b of dddt_b
      double Fb, dFbr, dFbh;
      wendF(C.wcon, r, in.h[b], Fb, dFbr, dFbh);
      const double ib = 1.0 / in.omega[b];
      ubar += gdb * ma * Fb * ib;
      k += gdb * ma * u * ib * dFbr;
    }
    // u = dv.dx: d/dvhat_a = dx, d/dr_a = dv; r = |dx|
    gvx += ubar * dx;
    gvy += ubar * dy;
    gvz += ubar * dz;
    const double kr = r > 0.0 ? k / r : 0.0;
    grx += ubar * dvx + kr * dx;
    gry += ubar * dvy + kr * dy;
    grz += ubar * dvz + kr * dz;
  }
  g.r[3 * a] += grx;
  g.r[3 * a + 1] += gry;
  g.r[3 * a + 2] += grz;
  g.vhat[3 * a] += gvx;
  g.vhat[3 * a + 1] += gvy;
  g.vhat[3 * a + 2] += gvz;
  g.h[a] += gha;
  g.omega[a] += -gia * ia * ia;
}

// ------------------------------------------------------ fluid-side coupling
// For fluid query a over its DEM partners d (the SPH->DEM coupling set),
// W = W(r, Rc/2) (calcPorosityIncr sph.h:1246-1256):
//   por_sum_a = sum_d POROSITY_INCR W                   (calcPorosity)
//   shep_d   += (m_a / rho_a) W                         (calcShepSum)
//   drag_a    = -sum_d m_d W fdrag_d / (rho_a shep_d)   (calcDEMtoSPHDrag)
// sph.h:1303-1362.  The Shepard sums need every fluid partner of d before
// the drag, so the forward is two launches, and so is the backward (the
// adjoint of shep_d is complete only after the first).
__global__ void __launch_bounds__(kTPB) k_cpl_fwd1(int32_t nq, const int32_t* __restrict__ query,
                           const int64_t* __restrict__ off, const int32_t* __restrict__ pd,
                           const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                           sphdem_adj_consts C, double* __restrict__ por_sum,
                           double* __restrict__ shep) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double3 ra = ld3(in.r, a);
  const double va = in.mass[a] / in.dens[a], hc = 0.5 * C.coupling_radius;
  double por = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int d = pd[j];
    const double3 rd = imageOf(in.r, d, pimg[j], C.shift_len);
    const double dx = ra.x - rd.x, dy = ra.y - rd.y, dz = ra.z - rd.z;
    double W, t0, t1;
    wendW(C.wcon, sqrt(dx * dx + dy * dy + dz * dz), hc, W, t0, t1);
    por += C.porosity_incr * W;
    if (W != 0.0) add1(shep, d, va * W);
  }
  por_sum[a] += por;
}

__global__ void __launch_bounds__(kTPB) k_cpl_fwd2(int32_t nq, const int32_t* __restrict__ query,
                           const int64_t* __restrict__ off, const int32_t* __restrict__ pd,
                           const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                           sphdem_adj_consts C, double* __restrict__ drag) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double3 ra = ld3(in.r, a);
  const double rhoa = in.dens[a], hc = 0.5 * C.coupling_radius;
  double x = 0.0, y = 0.0, z = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int d = pd[j];
    const double sd = in.shep[d];
    if (!(sd > 0.0)) continue;
    const double3 rd = imageOf(in.r, d, pimg[j], C.shift_len);
    const double dx = ra.x - rd.x, dy = ra.y - rd.y, dz = ra.z - rd.z;
    double W, t0, t1;
    wendW(C.wcon, sqrt(dx * dx + dy * dy + dz * dz), hc, W, t0, t1);
    const double c = in.mass[d] * W / (rhoa * sd);
    const double3 fd = ld3(in.fdrag, d);
    x -= c * fd.x;
    y -= c * fd.y;
    z -= c * fd.z;
  }
  drag[3 * a] += x;
  drag[3 * a + 1] += y;
  drag[3 * a + 2] += z;
}

// first backward launch: the drag and porosity terms; shepbar[d] collects the
// adjoint of the Shepard sums
__global__ void __launch_bounds__(kTPB) k_cpl_bwd1(int32_t nq, const int32_t* __restrict__ query,
                           const int64_t* __restrict__ off, const int32_t* __restrict__ pd,
                           const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                           sphdem_adj_consts C, const double* __restrict__ gpor,
                           const double* __restrict__ gdrag, double* __restrict__ shepbar,
                           sphdem_adj_grads g) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double3 ra = ld3(in.r, a), gdr = ld3(gdrag, a);
  const double gp = gpor[a];
  const double rhoa = in.dens[a], hc = 0.5 * C.coupling_radius;
  double grx = 0.0, gry = 0.0, grz = 0.0, gda = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int d = pd[j];
    const double3 rd = imageOf(in.r, d, pimg[j], C.shift_len);
    const double dx = ra.x - rd.x, dy = ra.y - rd.y, dz = ra.z - rd.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double W, dWr, t1;
    wendW(C.wcon, r, hc, W, dWr, t1);
    double Wbar = gp * C.porosity_incr;
    const double sd = in.shep[d];
    if (sd > 0.0) {
      const double md = in.mass[d];
      const double c = md * W / (rhoa * sd);
      const double3 fd = ld3(in.fdrag, d);
      const double cbar = -(gdr.x * fd.x + gdr.y * fd.y + gdr.z * fd.z);
      add3(g.fdrag, d, -c * gdr.x, -c * gdr.y, -c * gdr.z);
      if (cbar != 0.0) add1(shepbar, d, -cbar * c / sd);
      gda += -cbar * c / rhoa;
      Wbar += cbar * md / (rhoa * sd);
    }
    if (r > 0.0 && Wbar != 0.0) {
      const double k = Wbar * dWr / r;
      grx += k * dx;
      gry += k * dy;
      grz += k * dz;
      add3(g.r, d, -k * dx, -k * dy, -k * dz);
    }
  }
  add3(g.r, a, grx, gry, grz);
  add1(g.dens, a, gda);
}

// second backward launch: the Shepard sums
__global__ void __launch_bounds__(kTPB) k_cpl_bwd2(int32_t nq, const int32_t* __restrict__ query,
                           const int64_t* __restrict__ off, const int32_t* __restrict__ pd,
                           const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                           sphdem_adj_consts C, const double* __restrict__ shepbar,
                           sphdem_adj_grads g) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int a = query[i];
  const double3 ra = ld3(in.r, a);
  const double rhoa = in.dens[a], va = in.mass[a] / rhoa, hc = 0.5 * C.coupling_radius;
  double grx = 0.0, gry = 0.0, grz = 0.0, gda = 0.0;
  for (int64_t j = off[i]; j < off[i + 1]; j++) {
    const int d = pd[j];
    const double sb = shepbar[d];
    if (sb == 0.0) continue;
    const double3 rd = imageOf(in.r, d, pimg[j], C.shift_len);
    const double dx = ra.x - rd.x, dy = ra.y - rd.y, dz = ra.z - rd.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double W, dWr, t1;
    wendW(C.wcon, r, hc, W, dWr, t1);
    // shep_d += (m_a / rho_a) W
    gda += -sb * va * W / rhoa;
    if (r > 0.0) {
      const double k = sb * va * dWr / r;
      grx += k * dx;
      gry += k * dy;
      grz += k * dz;
      add3(g.r, d, -k * dx, -k * dy, -k * dz);
    }
  }
  add3(g.r, a, grx, gry, grz);
  add1(g.dens, a, gda);
}

// ------------------------------------------------ fluid-particle interaction
// For DEM particle p over fluid partners n (W = W(|r_p - r_n'|, h_n),
// dvol = m_n / rho_n):
//   S  = sum dvol W,  U = sum dvol W vhat_n,  Pn = sum dvol W por_n,
//   Gp = sum m_n W fpfb_n,  Gv = sum (m_n W + dvol W) fv_n
// and, when S > 0.5, with fvel = U/S, por = Pn/S, fpfb = Gp/S, fvv = Gv/S:
//   dv = por (fvel - vhat_p),  vdiff = |dv|,  Re = (2R/nu) vdiff,
//   chi = drag_a - drag_b exp(-(1.5 - log10 Re)^2 / 2),
//   C = (0.63 + 4.8 / sqrt(Re))^2,  s = A C vdiff / 2  (A = pi R^2),
//   fd = s por^(-chi) dv  (0 when vdiff = 0)
//   fdrag_p = (V/m_p)(fvv + fpfb) + (rho0/m_p) fd
struct Fpi {
  double S, Pn;
  double3 U, Gp, Gv;
};

__device__ __forceinline__ void fpiSums(const sphdem_adj_fields& in, const sphdem_adj_consts& C,
                                        double3 rp, int64_t j0, int64_t j1,
                                        const int32_t* __restrict__ pb,
                                        const int32_t* __restrict__ pimg, Fpi& o) {
  o.S = o.Pn = 0.0;
  o.U = o.Gp = o.Gv = make_double3(0.0, 0.0, 0.0);
  for (int64_t j = j0; j < j1; j++) {
    const int n = pb[j];
    const double3 rn = imageOf(in.r, n, pimg[j], C.shift_len);
    const double dx = rp.x - rn.x, dy = rp.y - rn.y, dz = rp.z - rn.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double W, t0, t1;
    wendW(C.wcon, r, in.h[n], W, t0, t1);
    const double mn = in.mass[n];
    const double dvW = (mn / in.dens[n]) * W, mW = mn * W;
    const double3 vn = ld3(in.vhat, n), fpn = ld3(in.fpfb, n), fvn = ld3(in.fv, n);
    o.S += dvW;
    o.Pn += dvW * in.por[n];
    o.U.x += dvW * vn.x;
    o.U.y += dvW * vn.y;
    o.U.z += dvW * vn.z;
    o.Gp.x += mW * fpn.x;
    o.Gp.y += mW * fpn.y;
    o.Gp.z += mW * fpn.z;
    o.Gv.x += (mW + dvW) * fvn.x;
    o.Gv.y += (mW + dvW) * fvn.y;
    o.Gv.z += (mW + dvW) * fvn.z;
  }
}

__global__ void __launch_bounds__(kTPB) k_fpi_fwd(int32_t nq, const int32_t* __restrict__ query,
                          const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                          const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                          sphdem_adj_consts C, double* __restrict__ fdrag) {
  const int i = xcdIndex();
  if (i >= nq) return;
  const int p = query[i];
  const double3 rp = ld3(in.r, p);
  Fpi o;
  fpiSums(in, C, rp, off[i], off[i + 1], pb, pimg, o);
  if (!(o.S > 0.5)) return;
  const double iS = 1.0 / o.S;
  const double por = o.Pn * iS;
  const double3 vp = ld3(in.vhat, p);
  const double dvx = por * (o.U.x * iS - vp.x), dvy = por * (o.U.y * iS - vp.y),
               dvz = por * (o.U.z * iS - vp.z);
  const double vdiff = sqrt(dvx * dvx + dvy * dvy + dvz * dvz);
  double sp = 0.0;
  if (vdiff != 0.0) {
    const double R = C.dem_radius;
    const double Re = (2.0 * R / C.viscosity) * vdiff;
    const double L10 = log10(Re);
    const double chi = C.drag_a - C.drag_b * exp(-(1.5 - L10) * (1.5 - L10) / 2.0);
    const double cc = 0.63 + 4.8 / sqrt(Re);
    const double s = 0.5 * (C.pi * R * R) * (cc * cc) * vdiff;
    sp = s * pow(por, -chi);
  }
  const double mp = in.mass[p];
  const double kv = C.dem_vol / mp, kd = C.dens / mp;
  fdrag[3 * p] += kv * (o.Gv.x + o.Gp.x) * iS + kd * sp * dvx;
  fdrag[3 * p + 1] += kv * (o.Gv.y + o.Gp.y) * iS + kd * sp * dvy;
  fdrag[3 * p + 2] += kv * (o.Gv.z + o.Gp.z) * iS + kd * sp * dvz;
}

__global__ void __launch_bounds__(kTPB) k_fpi_bwd(int32_t nq, const int32_t* __restrict__ query,
                          const int64_t* __restrict__ off, const int32_t* __restrict__ pb,
                          const int32_t* __restrict__ pimg, sphdem_adj_fields in,
                          sphdem_adj_consts C, const double* __restrict__ gfd,
                          sphdem_adj_grads g) {
  const int i = xcdIndex();
  double gpar[3] = {0.0, 0.0, 0.0};  // nu, drag_a, drag_b
  const int p = i < nq ? query[i] : 0;
  Fpi o;
  const double3 rp = i < nq ? ld3(in.r, p) : make_double3(0.0, 0.0, 0.0);
  if (i < nq) fpiSums(in, C, rp, off[i], off[i + 1], pb, pimg, o);
  if (i < nq && o.S > 0.5) {
    const double iS = 1.0 / o.S;
    const double por = o.Pn * iS;
    const double3 fvel = make_double3(o.U.x * iS, o.U.y * iS, o.U.z * iS);
    const double3 vp = ld3(in.vhat, p);
    const double ex = fvel.x - vp.x, ey = fvel.y - vp.y, ez = fvel.z - vp.z;
    const double dvx = por * ex, dvy = por * ey, dvz = por * ez;
    const double vdiff = sqrt(dvx * dvx + dvy * dvy + dvz * dvz);
    const double mp = in.mass[p];
    const double kv = C.dem_vol / mp, kd = C.dens / mp;
    const double3 gf = ld3(gfd, p);
    // fdrag = kv (Gv + Gp)/S + kd fd
    const double3 Gbar = make_double3(kv * gf.x * iS, kv * gf.y * iS, kv * gf.z * iS);
    double Sbar = -(Gbar.x * (o.Gv.x + o.Gp.x) + Gbar.y * (o.Gv.y + o.Gp.y) +
                    Gbar.z * (o.Gv.z + o.Gp.z)) * iS;
    double porbar = 0.0;
    double3 fvelbar = make_double3(0.0, 0.0, 0.0);
    if (vdiff != 0.0) {
      const double R = C.dem_radius, A = C.pi * R * R;
      const double Re = (2.0 * R / C.viscosity) * vdiff;
      const double L10 = log10(Re);
      const double E = exp(-(1.5 - L10) * (1.5 - L10) / 2.0);
      const double chi = C.drag_a - C.drag_b * E;
      const double cc = 0.63 + 4.8 / sqrt(Re);
      const double Cd = cc * cc;
      const double s = 0.5 * A * Cd * vdiff;
      const double pw = pow(por, -chi);
      // fd = (s pw) dv
      const double fdx = kd * gf.x, fdy = kd * gf.y, fdz = kd * gf.z;
      const double tau = fdx * dvx + fdy * dvy + fdz * dvz;
      double dvbx = s * pw * fdx, dvby = s * pw * fdy, dvbz = s * pw * fdz;
      const double sbar = tau * pw, pwbar = tau * s;
      // pw = por^(-chi)
      const double chibar = -log(por) * pw * pwbar;
      porbar += -chi * pw / por * pwbar;
      gpar[1] = chibar;
      gpar[2] = -E * chibar;
      const double Ebar = -C.drag_b * chibar;
      // s = A Cd vdiff / 2;  Cd = (0.63 + 4.8 Re^-1/2)^2;  E = exp(-(1.5 - L10)^2/2)
      double vdbar = sbar * 0.5 * A * Cd;
      const double Cdbar = sbar * 0.5 * A * vdiff;
      const double dCd = -4.8 * cc * pow(Re, -1.5);
      const double dE = E * (1.5 - L10) / (Re * log(10.0));
      const double Rebar = Cdbar * dCd + Ebar * dE;
      // Re = (2R/nu) vdiff
      vdbar += Rebar * (2.0 * R / C.viscosity);
      gpar[0] = -Rebar * Re / C.viscosity;
      // vdiff = |dv|
      dvbx += vdbar * dvx / vdiff;
      dvby += vdbar * dvy / vdiff;
      dvbz += vdbar * dvz / vdiff;
      // dv = por (fvel - vhat_p)
      porbar += dvbx * ex + dvby * ey + dvbz * ez;
      fvelbar = make_double3(por * dvbx, por * dvby, por * dvbz);
      add3(g.vhat, p, -por * dvbx, -por * dvby, -por * dvbz);
    }
    // fvel = U/S, por = Pn/S
    const double3 Ubar = make_double3(fvelbar.x * iS, fvelbar.y * iS, fvelbar.z * iS);
    Sbar -= (fvelbar.x * fvel.x + fvelbar.y * fvel.y + fvelbar.z * fvel.z) * iS;
    const double Pnbar = porbar * iS;
    Sbar -= porbar * por * iS;
    // back through the sums, pair by pair
    double rpx = 0.0, rpy = 0.0, rpz = 0.0;
    for (int64_t j = off[i]; j < off[i + 1]; j++) {
      const int n = pb[j];
      const double3 rn = imageOf(in.r, n, pimg[j], C.shift_len);
      const double dx = rp.x - rn.x, dy = rp.y - rn.y, dz = rp.z - rn.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      double W, dWdr, dWdh;
      wendW(C.wcon, r, in.h[n], W, dWdr, dWdh);
      if (W == 0.0 && dWdr == 0.0 && dWdh == 0.0) continue;
      const double mn = in.mass[n], rhon = in.dens[n];
      const double dvol = mn / rhon;
      const double dvW = dvol * W, mW = mn * W;
      const double3 vn = ld3(in.vhat, n), fpn = ld3(in.fpfb, n), fvn = ld3(in.fv, n);
      const double porn = in.por[n];
      const double dvWbar = Ubar.x * vn.x + Ubar.y * vn.y + Ubar.z * vn.z + Pnbar * porn + Sbar +
                            Gbar.x * fvn.x + Gbar.y * fvn.y + Gbar.z * fvn.z;
      const double mWbar = Gbar.x * (fpn.x + fvn.x) + Gbar.y * (fpn.y + fvn.y) +
                           Gbar.z * (fpn.z + fvn.z);
      add3(g.vhat, n, dvW * Ubar.x, dvW * Ubar.y, dvW * Ubar.z);
      add1(g.por, n, dvW * Pnbar);
      add3(g.fpfb, n, mW * Gbar.x, mW * Gbar.y, mW * Gbar.z);
      add3(g.fv, n, (mW + dvW) * Gbar.x, (mW + dvW) * Gbar.y, (mW + dvW) * Gbar.z);
      const double Wbar = dvWbar * dvol + mWbar * mn;
      add1(g.dens, n, -dvWbar * W * mn / (rhon * rhon));
      add1(g.h, n, Wbar * dWdh);
      if (r > 0.0) {
        const double c = Wbar * dWdr / r;
        rpx += c * dx;
        rpy += c * dy;
        rpz += c * dz;
        add3(g.r, n, -c * dx, -c * dy, -c * dz);
      }
    }
    add3(g.r, p, rpx, rpy, rpz);
  }
  blockAdd<3>(gpar, g.params);
}

// -------------------------------------------------------------- DEM contacts
// Pair (a, b): dx = r_a - r_b', r = |dx|, n = dx/r, overlap = 2R - r; when
// overlap > 0: od = -(vhat_a - vhat_b).n, sN = k overlap + gamma od,
// fc_a += sN n / m_a.
__global__ void __launch_bounds__(kTPB) k_contacts_fwd(int64_t np, const int32_t* __restrict__ pa,
                               const int32_t* __restrict__ pb, const int32_t* __restrict__ pimg,
                               sphdem_adj_fields in, sphdem_adj_consts C,
                               double* __restrict__ fc) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int a = pa[k], b = pb[k];
  const double3 ra = ld3(in.r, a), rb = imageOf(in.r, b, pimg[k], C.shift_len);
  const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double overlap = 2.0 * C.dem_radius - r;
  if (!(overlap > 0.0) || !(r > 0.0)) return;
  const double nx = dx / r, ny = dy / r, nz = dz / r;
  const double3 va = ld3(in.vhat, a), vb = ld3(in.vhat, b);
  const double od = -((va.x - vb.x) * nx + (va.y - vb.y) * ny + (va.z - vb.z) * nz);
  const double s = (C.dem_k * overlap + C.dem_gamma * od) / in.mass[a];
  add3(fc, a, s * nx, s * ny, s * nz);
}

__global__ void __launch_bounds__(kTPB) k_contacts_bwd(int64_t np, const int32_t* __restrict__ pa,
                               const int32_t* __restrict__ pb, const int32_t* __restrict__ pimg,
                               sphdem_adj_fields in, sphdem_adj_consts C,
                               const double* __restrict__ gfc, sphdem_adj_grads g) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double gpar[2] = {0.0, 0.0};  // dem_k, dem_gamma
  if (k < np) {
    const int a = pa[k], b = pb[k];
    const double3 ra = ld3(in.r, a), rb = imageOf(in.r, b, pimg[k], C.shift_len);
    const double dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double overlap = 2.0 * C.dem_radius - r;
    if (overlap > 0.0 && r > 0.0) {
      const double nx = dx / r, ny = dy / r, nz = dz / r;
      const double3 va = ld3(in.vhat, a), vb = ld3(in.vhat, b);
      const double dvx = va.x - vb.x, dvy = va.y - vb.y, dvz = va.z - vb.z;
      const double od = -(dvx * nx + dvy * ny + dvz * nz);
      const double im = 1.0 / in.mass[a];
      const double sN = C.dem_k * overlap + C.dem_gamma * od;
      const double3 gf = ld3(gfc, a);
      // fc = sN n / m
      const double sNbar = (gf.x * nx + gf.y * ny + gf.z * nz) * im;
      double nbx = sN * im * gf.x, nby = sN * im * gf.y, nbz = sN * im * gf.z;
      gpar[0] = sNbar * overlap;
      gpar[1] = sNbar * od;
      const double odbar = C.dem_gamma * sNbar;
      double rbar = -C.dem_k * sNbar;  // overlap = 2R - r
      // od = -(dv . n)
      add3(g.vhat, a, -odbar * nx, -odbar * ny, -odbar * nz);
      add3(g.vhat, b, odbar * nx, odbar * ny, odbar * nz);
      nbx -= odbar * dvx;
      nby -= odbar * dvy;
      nbz -= odbar * dvz;
      // n = dx / r: d n / d dx = (I - n n^T) / r;  r = |dx|
      const double nn = nbx * nx + nby * ny + nbz * nz;
      const double gx = (nbx - nn * nx) / r + rbar * nx;
      const double gy = (nby - nn * ny) / r + rbar * ny;
      const double gz = (nbz - nn * nz) / r + rbar * nz;
      add3(g.r, a, gx, gy, gz);
      add3(g.r, b, -gx, -gy, -gz);
    }
  }
  blockAdd<2>(gpar, g.params + 3);
}

inline unsigned gridFor(int64_t n) { return (unsigned)((n + kTPB - 1) / kTPB); }

}  // namespace sphdem_adj

// ---------------------------------------------------------------- C ABI
namespace {
int adjLaunched() { return hipGetLastError() == hipSuccess ? 0 : SPHDEM_E_HIP; }
}  // namespace

// ---- device-resident scratch of sphdem_adj_coupling_bwd_noalloc
// The Shepard-sum adjoints of the coupling backward (one double per particle),
// for callers that cannot allocate.  Every GPU the library runs on holds its
// own copy for the life of the process; host code reaches it only through the
// per-device address hipGetSymbolAddress returns (the symbol itself is a
// host-side stand-in, not a device address).
static __device__ __align__(256) double scratch[SPHDEM_ADJ_NOALLOC_MAX_PARTICLES];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratchMu;
double* g_scratchPtr[kScratchDevices] = {};

// The current device's scratch address, looked up once per device (with that
// device current) on the first call made on it.
int scratchOfCurrentDevice(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) return SPHDEM_E_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SPHDEM_E_HIP;
    g_scratchPtr[dev] = (double*)p;
  }
  *out = g_scratchPtr[dev];
  return 0;
}

// zero fill (a kernel, not a memset: the call stays a plain kernel sequence)
__global__ void __launch_bounds__(sphdem_adj::kTPB) k_zero_doubles(double* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
}  // namespace

extern "C" {

#define SPHDEM_ADJ_CSR_OK(in)                                                       \
  (query && off && b && img && (in) && (in)->r && (in)->mass)

int sphdem_adj_force(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                     const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                     double* fv, double* fp, void* stream) {
  if (nquery < 0 || !in || !c || !fv || !fp) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->h || !in->vhat || !in->dens || !in->omega || !in->pdr2 ||
      !in->iam)
    return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_force_fwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, fv, fp);
  return adjLaunched();
}

int sphdem_adj_force_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                         const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                         const sphdem_adj_consts* c, const double* gfv, const double* gfp,
                         const sphdem_adj_grads* g, void* stream) {
  if (nquery < 0 || !in || !c || !gfv || !gfp || !g) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->h || !in->vhat || !in->dens || !in->omega || !in->pdr2 ||
      !in->iam || !g->r || !g->vhat || !g->h || !g->dens || !g->omega || !g->pdr2 || !g->params)
    return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_force_bwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, gfv, gfp, *g);
  return adjLaunched();
}

int sphdem_adj_omega(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                     const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                     double* omega, void* stream) {
  if (nquery < 0 || !in || !c || !omega) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->h || !in->dens) return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_omega_fwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, omega);
  return adjLaunched();
}

int sphdem_adj_omega_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                         const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                         const sphdem_adj_consts* c, const double* gomega,
                         const sphdem_adj_grads* g, void* stream) {
  if (nquery < 0 || !in || !c || !gomega || !g) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->h || !in->dens || !g->r || !g->h || !g->dens)
    return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_omega_bwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, gomega, *g);
  return adjLaunched();
}

int sphdem_adj_dddt(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                    const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                    double* dddt, void* stream) {
  if (nquery < 0 || !in || !c || !dddt) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->h || !in->vhat || !in->omega) return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_dddt_fwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, dddt);
  return adjLaunched();
}

int sphdem_adj_dddt_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                        const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                        const sphdem_adj_consts* c, const double* gdddt,
                        const sphdem_adj_grads* g, void* stream) {
  if (nquery < 0 || !in || !c || !gdddt || !g) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->h || !in->vhat || !in->omega || !g->r || !g->vhat || !g->h ||
      !g->omega)
    return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_dddt_bwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, gdddt, *g);
  return adjLaunched();
}

int sphdem_adj_coupling(int32_t nquery, const int32_t* query, const int64_t* off,
                        const int32_t* d, const int32_t* img, const sphdem_adj_fields* in,
                        const sphdem_adj_consts* c, double* por_sum, double* shep, double* drag,
                        void* stream) {
  if (nquery < 0 || !in || !c || !por_sum || !shep || !drag) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  const int32_t* b = d;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->dens || !in->fdrag) return SPHDEM_E_ARG;
  const dim3 grid(sphdem_adj::gridFor(nquery)), block(sphdem_adj::kTPB);
  hipLaunchKernelGGL(sphdem_adj::k_cpl_fwd1, grid, block, 0, (hipStream_t)stream, nquery, query,
                     off, d, img, *in, *c, por_sum, shep);
  if (adjLaunched()) return SPHDEM_E_HIP;
  sphdem_adj_fields in2 = *in;
  in2.shep = shep;
  hipLaunchKernelGGL(sphdem_adj::k_cpl_fwd2, grid, block, 0, (hipStream_t)stream, nquery, query,
                     off, d, img, in2, *c, drag);
  return adjLaunched();
}

int sphdem_adj_coupling_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                            const int32_t* d, const int32_t* img, const sphdem_adj_fields* in,
                            const sphdem_adj_consts* c, const double* gpor, const double* gdrag,
                            double* work, const sphdem_adj_grads* g, void* stream) {
  if (nquery < 0 || !in || !c || !gpor || !gdrag || !work || !g) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  const int32_t* b = d;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->dens || !in->fdrag || !in->shep || !g->r || !g->dens ||
      !g->fdrag)
    return SPHDEM_E_ARG;
  const dim3 grid(sphdem_adj::gridFor(nquery)), block(sphdem_adj::kTPB);
  hipLaunchKernelGGL(sphdem_adj::k_cpl_bwd1, grid, block, 0, (hipStream_t)stream, nquery, query,
                     off, d, img, *in, *c, gpor, gdrag, work, *g);
  if (adjLaunched()) return SPHDEM_E_HIP;
  hipLaunchKernelGGL(sphdem_adj::k_cpl_bwd2, grid, block, 0, (hipStream_t)stream, nquery, query,
                     off, d, img, *in, *c, (const double*)work, *g);
  return adjLaunched();
}

int sphdem_adj_coupling_bwd_noalloc(int32_t nquery, const int32_t* query, const int64_t* off,
                                    const int32_t* d, const int32_t* img,
                                    const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                                    const double* gpor, const double* gdrag, int64_t n,
                                    const sphdem_adj_grads* g, void* stream) {
  if (nquery < 0 || n < 0 || !in || !c || !gpor || !gdrag || !g) return SPHDEM_E_ARG;
  if (n > SPHDEM_ADJ_NOALLOC_MAX_PARTICLES) return SPHDEM_E_CAPACITY;
  if (nquery == 0) return 0;
  const int32_t* b = d;
  if (!SPHDEM_ADJ_CSR_OK(in) || !in->dens || !in->fdrag || !in->shep || !g->r || !g->dens ||
      !g->fdrag)
    return SPHDEM_E_ARG;
  double* work = nullptr;
  const int rc = scratchOfCurrentDevice(&work);
  if (rc) return rc;
  // the array keeps the last call's contents: clear the part this call uses
  if (n > 0) {
    hipLaunchKernelGGL(k_zero_doubles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, work, n);
    if (adjLaunched()) return SPHDEM_E_HIP;
  }
  const dim3 grid(sphdem_adj::gridFor(nquery)), block(sphdem_adj::kTPB);
  hipLaunchKernelGGL(sphdem_adj::k_cpl_bwd1, grid, block, 0, (hipStream_t)stream, nquery, query,
                     off, d, img, *in, *c, gpor, gdrag, work, *g);
  if (adjLaunched()) return SPHDEM_E_HIP;
  hipLaunchKernelGGL(sphdem_adj::k_cpl_bwd2, grid, block, 0, (hipStream_t)stream, nquery, query,
                     off, d, img, *in, *c, (const double*)work, *g);
  return adjLaunched();
}

int sphdem_adj_fpi(int32_t nquery, const int32_t* query, const int64_t* off, const int32_t* b,
                   const int32_t* img, const sphdem_adj_fields* in, const sphdem_adj_consts* c,
                   double* fdrag, void* stream) {
  if (nquery < 0 || !in || !c || !fdrag) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!query || !off || !b || !img || !in->r || !in->vhat || !in->h || !in->dens ||
      !in->mass || !in->fpfb || !in->fv || !in->por)
    return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_fpi_fwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, fdrag);
  return adjLaunched();
}

int sphdem_adj_fpi_bwd(int32_t nquery, const int32_t* query, const int64_t* off,
                       const int32_t* b, const int32_t* img, const sphdem_adj_fields* in,
                       const sphdem_adj_consts* c, const double* gfdrag,
                       const sphdem_adj_grads* g, void* stream) {
  if (nquery < 0 || !in || !c || !gfdrag || !g) return SPHDEM_E_ARG;
  if (nquery == 0) return 0;
  if (!query || !off || !b || !img || !in->r || !in->vhat || !in->h || !in->dens ||
      !in->mass || !in->fpfb || !in->fv || !in->por || !g->r || !g->vhat || !g->h ||
      !g->dens || !g->fpfb || !g->fv || !g->por || !g->params)
    return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_fpi_bwd, dim3(sphdem_adj::gridFor(nquery)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, nquery, query, off, b, img,
                     *in, *c, gfdrag, *g);
  return adjLaunched();
}

int sphdem_adj_contacts(int64_t npairs, const int32_t* a, const int32_t* b, const int32_t* img,
                        const sphdem_adj_fields* in, const sphdem_adj_consts* c, double* fc,
                        void* stream) {
  if (npairs < 0 || !in || !c || !fc) return SPHDEM_E_ARG;
  if (npairs == 0) return 0;
  if (!a || !b || !img || !in->r || !in->vhat || !in->mass) return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_contacts_fwd, dim3(sphdem_adj::gridFor(npairs)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, npairs, a, b, img, *in, *c,
                     fc);
  return adjLaunched();
}

int sphdem_adj_contacts_bwd(int64_t npairs, const int32_t* a, const int32_t* b,
                            const int32_t* img, const sphdem_adj_fields* in,
                            const sphdem_adj_consts* c, const double* gfc,
                            const sphdem_adj_grads* g, void* stream) {
  if (npairs < 0 || !in || !c || !gfc || !g) return SPHDEM_E_ARG;
  if (npairs == 0) return 0;
  if (!a || !b || !img || !in->r || !in->vhat || !in->mass || !g->r || !g->vhat || !g->params)
    return SPHDEM_E_ARG;
  hipLaunchKernelGGL(sphdem_adj::k_contacts_bwd, dim3(sphdem_adj::gridFor(npairs)),
                     dim3(sphdem_adj::kTPB), 0, (hipStream_t)stream, npairs, a, b, img, *in, *c,
                     gfc, *g);
  return adjLaunched();
}

}  // extern "C"
