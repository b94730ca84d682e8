// sphdem_mixed.hip -- the mixed-precision form of the SPH pair passes.
//
// Everything that decides which pairs exist, and everything the integrator
// carries, stays FP64: the cached neighbour lists (built by the FP64 search,
// so the pair sets are the reference's exactly), positions, h, the per-particle
// sums (each pair's contribution is accumulated in FP64) and the dt
// reduction.  What the pair passes gather of a neighbour is a compact 16-byte
// record per field group, packed from the FP64 store right before the pass:
//   P: position as 32-bit fixed point over the cell grid (unit a power of two,
//      grid extent / 2^32 or finer), mass as FP32
//   V: vhat and h as FP32
//   D: dens, pdr2, 1/h, 1/h^5 as FP32
// so a neighbour costs one to three 16-byte loads instead of three to six
// 32-byte ones.  Position differences are exact in the fixed-point units and
// rounded once to FP32; the per-pair kernel and force arithmetic runs in FP32
// with hardware reciprocal and square root and the factored Wendland forms
//   F(q)    = -10 WCON (2-q)^3 / h^5             (the reference divides
//   dK/dh   =  10 WCON q^2 (2-q)^3 / h             -4(2-q)^3(1+2q)+2(2-q)^4 = -10q(2-q)^3
//                                                  by q, cancelling near q = 0)
// which are algebraically the reference's (sph.h:74-81, 285-292) without its
// cancellation near q = 0.  Errors against the FP64 form are measured by the
// tests and the benchmark.
#pragma once

namespace sphdem {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float cube_f(float x) { return x * x * x; }

// F(q)/h^5 factor: -10 WCON (2-q)^3 i5, zero beyond the support
__device__ __forceinline__ float wendF_m(float wcon10, float q, float i5) {
  const float t = 2.0f - q;
  return q <= 2.0f ? -wcon10 * cube_f(t) * i5 : 0.0f;
}
// K(q) = WCON (2-q)^4 (1+2q)
__device__ __forceinline__ float wendK_m(float wcon, float q) {
  const float t = 2.0f - q;
  return q <= 2.0f ? wcon * (t * t) * (t * t) * (1.0f + 2.0f * q) : 0.0f;
}
// dW/dh = (1/h^3) dK/dh - 3/h^4 K, dK/dh = 10 WCON q^2 (2-q)^3 / h
__device__ __forceinline__ float wendPdWh_m(float wcon, float q, float rh, float i3, float c3) {
  const float t = 2.0f - q;
  const float pdKh = q <= 2.0f ? 10.0f * wcon * (q * q) * cube_f(t) * rh : 0.0f;
  return i3 * pdKh - c3 * wendK_m(wcon, q);
}

// (Quant, quant1, mixRecordP and mixRecordV: sphdem_kernels.h)
struct DxF {
  float x, y, z, r2;
};
// a - b from packed positions: exact in fixed point, then one rounding
__device__ __forceinline__ DxF dxq(const uint4& a, const uint4& b, float res) {
  DxF d;
  d.x = (float)(int)(a.x - b.x) * res;
  d.y = (float)(int)(a.y - b.y) * res;
  d.z = (float)(int)(a.z - b.z) * res;
  d.r2 = d.x * d.x + d.y * d.y + d.z * d.z;
  return d;
}

// Packs the records of the SPH-group entries [0, ns) and the ghosts
// [n, n + ng); groups: bit 0 P, bit 1 V, bit 2 D.
constexpr unsigned MIX_P = 1, MIX_V = 2, MIX_D = 4;
__global__ void k_pack_mix(Store S, int ns, int n, int ng, Quant Q, unsigned groups,
                           uint4* __restrict__ P, float4* __restrict__ V, float4* __restrict__ D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns + ng) return;
  const int e = i < ns ? i : n + (i - ns);
  const double4 p = S.posh[e];
  if (groups & MIX_P) P[e] = mixRecordP(p, S.vhm[e], Q);
  if (groups & MIX_V) V[e] = mixRecordV(p, S.vhm[e]);
  if (groups & MIX_D) {
    const double4 d = S.dpo[e];
    const double rh = 1.0 / p.w;
    D[e] = make_float4((float)d.x, (float)d.y, (float)rh, (float)(rh * rh * rh * rh * rh));
  }
}

// ---- calcOmega, mixed --------------------------------------------------------------
__global__ void __launch_bounds__(256) k_omega_m(Store S, int ns, const int* __restrict__ nbl,
                                                 const int* __restrict__ nbcnt,
                                                 const uint4* __restrict__ P, float res,
                                                 double wcon) {
  const int a = xcdThread();
  if (a >= ns) return;
  const uint4 qa = P[a];
  const double ha = S.posh[a].w;
  const float wc = (float)wcon;
  const float rha = (float)(1.0 / ha), i3 = (float)(1.0 / cube(ha)), c3 = (float)(3.0 / pow4(ha));
  double om = S.vhm[a].w * (double)wendPdWh_m(wc, 0.0f, rha, i3, c3);
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    uint4 q[U];
#pragma unroll
    for (int u = 0; u < U; u++) q[u] = P[e[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      const DxF d = dxq(qa, q[u], res);
      const float qq = fsqrt(d.r2) * rha;
      om += (double)__uint_as_float(q[u].w) * (double)wendPdWh_m(wc, qq, rha, i3, c3);
    }
  }
  S.dpo[a].z = 1 + ha * om / (S.dpo[a].x * 3);
}

// ---- calcDensity, mixed ------------------------------------------------------------
__global__ void __launch_bounds__(256) k_density_m(Store S, int ns, const int* __restrict__ nbl,
                                                   const int* __restrict__ nbcnt,
                                                   const uint4* __restrict__ P, float res,
                                                   double wcon) {
  const int a = xcdThread();
  if (a >= ns) return;
  const uint4 qa = P[a];
  const double ha = S.posh[a].w;
  const float wc = (float)wcon;
  const float rha = (float)(1.0 / ha);
  const double i3 = 1.0 / cube(ha);
  double rho = S.vhm[a].w * (i3 * (double)wendK_m(wc, 0.0f));
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    uint4 q[U];
#pragma unroll
    for (int u = 0; u < U; u++) q[u] = P[e[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      const DxF d = dxq(qa, q[u], res);
      const float qq = fsqrt(d.r2) * rha;
      rho += (double)__uint_as_float(q[u].w) * (i3 * (double)wendK_m(wc, qq));
    }
  }
  S.dpo[a].x = rho;
}

// ---- calcDddtDudt, mixed -----------------------------------------------------------
// OMG (VAR_H only): calcOmega fused into the walk (the k_omega_m sum, in list
// order); the dddt terms are summed without 1/omega_a and scaled once in FP64.
template <bool VARH, bool OMG = false>
__global__ void __launch_bounds__(256) k_dddt_m(Store S, int ns, const int* __restrict__ nbl,
                                                const int* __restrict__ nbcnt,
                                                const uint4* __restrict__ P,
                                                const float4* __restrict__ V, float res,
                                                double wcon) {
  static_assert(VARH || !OMG, "omega is only fused under VAR_H");
  const int a = xcdThread();
  if (a >= ns) return;
  const uint4 qa = P[a];
  const float4 va = V[a];
  const double ha = S.posh[a].w;
  const float wc10 = (float)(10.0 * wcon);
  const float rha = (float)(1.0 / ha), i5a = (float)(1.0 / pow5(ha));
  const float rom = (VARH && !OMG) ? (float)(1.0 / S.dpo[a].z) : 1.0f;
  const float wc = (float)wcon;
  const float i3 = OMG ? (float)(1.0 / cube(ha)) : 0.0f, c3 = OMG ? (float)(3.0 / pow4(ha)) : 0.0f;
  double om = OMG ? S.vhm[a].w * (double)wendPdWh_m(wc, 0.0f, rha, i3, c3) : 0.0;
  double dddt = OMG ? 0.0 : S.vh0d[a].w;
  const int n = nbcnt[a];
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += U) {
    SPHDEM_FETCH_ENTRIES(U)
    uint4 q[U];
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      q[u] = P[e[u]];
      v[u] = V[e[u]];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      const DxF d = dxq(qa, q[u], res);
      const float dv0 = va.x - v[u].x, dv1 = va.y - v[u].y, dv2 = va.z - v[u].z;
      const float r = fsqrt(d.r2);
      if (OMG) om += (double)__uint_as_float(q[u].w) * (double)wendPdWh_m(wc, r * rha, rha, i3, c3);
      float gF;  // gradW = dx * gF
      if (VARH) {
        gF = wendF_m(wc10, r * rha, i5a) * rom;
      } else {
        const float rhb = frcp(v[u].w);
        const float i5b = (rhb * rhb) * (rhb * rhb) * rhb;
        const float Fa = wendF_m(wc10, r * rha, i5a);
        const float Fb = wendF_m(wc10, r * rhb, i5b);
        gF = 0.5f * (Fa + Fb);
      }
      dddt += (double)__uint_as_float(q[u].w) * (double)(gF * (dv0 * d.x + dv1 * d.y + dv2 * d.z));
    }
  }
  if (OMG) {
    const double omega = 1 + ha * om / (S.dpo[a].x * 3);
    S.dpo[a].z = omega;
    dddt = S.vh0d[a].w + (1.0 / omega) * dddt;
  }
  S.vh0d[a].w = dddt;
}

// ---- calcForce, mixed --------------------------------------------------------------
// OMG (VAR_H only): calcOmega fused into the walk, as in k_dddt_m; the force
// sums are taken without 1/omega_a and scaled once in FP64.
template <bool VARH, int VISC, bool ARTIF, bool LIQDEM, bool OMG = false>
__global__ void __launch_bounds__(256) k_force_m(Store S, int ns, const int* __restrict__ nbl,
                                                 const int* __restrict__ nbcnt,
                                                 const uint4* __restrict__ Pq,
                                                 const float4* __restrict__ V,
                                                 const float4* __restrict__ D, float res,
                                                 sphdem_params P, DevGlobals* g) {
  static_assert(VARH || !OMG, "omega is only fused under VAR_H");
  const int a = xcdThread();
  const double inf = __longlong_as_double(0x7ff0000000000000ll);
  double newDt = inf;
  if (a < ns) {
    const bool fluid = S.iam[a] == T_SPH;
    const uint4 qa = Pq[a];
    const float4 va = V[a];
    const double4 da = S.dpo[a];
    const double ha = S.posh[a].w;
    const float wc10 = (float)(10.0 * P.wcon);
    const float rha = (float)(1.0 / ha), i5a = (float)(1.0 / pow5(ha));
    const float rom = (VARH && !OMG) ? (float)(1.0 / da.z) : 1.0f;
    const float wc = (float)P.wcon;
    const float i3 = OMG ? (float)(1.0 / cube(ha)) : 0.0f, c3 = OMG ? (float)(3.0 / pow4(ha)) : 0.0f;
    double om = OMG ? S.vhm[a].w * (double)wendPdWh_m(wc, 0.0f, rha, i3, c3) : 0.0;
    const float rhoa = (float)da.x, pdr2a = (float)da.y;
    const float nu = (float)P.viscosity, viscLiq = (float)(2.0 * P.viscosity * P.dens);
    const float cs2 = (float)(2.0 * P.spsound), alpha = (float)P.alpha_artificial;
    const float nu198 = (float)(19.8 * P.viscosity), alphaM = (float)monaghanAlpha(P);
    const double vsigC = (VISC == VISC_MOR || VISC == VISC_CLE) ? 2.0 * P.spsound : 0.0;
    const double cdiff = VISC == VISC_MON ? 0.06 : 0.125;  // viscDiffusionCondition
    // the dt conditions 0.6 hav and cdiff hav^2 are increasing in hav, and so
    // are their correctly rounded values: their minima are those of the
    // smallest hav, taken once at the end
    double minHav = inf;
    // f is the sum of the pressure, viscous and artificial-viscosity parts,
    // added once at the end
    double ax = 0.0, ay = 0.0, az = 0.0, px = 0.0, py = 0.0, pz = 0.0, vx = 0.0, vy = 0.0,
           vz = 0.0;
    const int n = nbcnt[a];
    constexpr int U = 8;
    for (int k0 = 0; k0 < n; k0 += U) {
      SPHDEM_FETCH_ENTRIES(U)
      uint4 q[U];
      float4 v[U], d[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        q[u] = Pq[e[u]];
        v[u] = V[e[u]];
        d[u] = D[e[u]];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (k0 + u >= n) break;
        const DxF dx = dxq(qa, q[u], res);
        const float mb = __uint_as_float(q[u].w), rhob = d[u].x, pdr2b = d[u].y;
        const float r = fsqrt(dx.r2);
        if (OMG) om += (double)mb * (double)wendPdWh_m(wc, r * rha, rha, i3, c3);
        const float Fa = wendF_m(wc10, r * rha, i5a);
        const float Fb = wendF_m(wc10, r * d[u].z, d[u].w);
        // gradients are dx * G; under VAR_H both divide by pa.omega (sph.h:331-332)
        float Ga, Gb;
        if (VARH) {
          Ga = Fa * rom;
          Gb = Fb * rom;
        } else {
          Ga = Gb = 0.5f * (Fa + Fb);
        }
        if (fluid) {
          const float dv0 = va.x - v[u].x, dv1 = va.y - v[u].y, dv2 = va.z - v[u].z;
          const float Gbar = 0.5f * (Ga + Gb);
          const double hav = 0.5 * (ha + (double)v[u].w);
          const float havf = (float)hav;
          float fvs = 0.0f;  // fv = dv * fvs (Morris)
          float fvr = 0.0f;  // fv = dx * fvr (Cleary, Monaghan)
          float vsig = (float)vsigC;
          const float vdr = (ARTIF || VISC == VISC_CLE || VISC == VISC_MON)
                                ? dx.x * dv0 + dx.y * dv1 + dx.z * dv2
                                : 0.0f;
          if (VISC == VISC_MOR) {
            const float rr = frcp(rhoa * rhob);
            const float visc = LIQDEM ? viscLiq * rr : nu * (rhoa + rhob) * rr;
            const float dd = Gbar * dx.r2;  // dot(dx, dx * Gbar)
            const float den = dx.r2 + 0.01f * (havf * havf);
            fvs = mb * visc * dd * frcp(den);
          }
          if (VISC == VISC_CLE) {
            const float den = dx.r2 + 0.01f * (havf * havf);
            fvr = mb * (nu198 * vdr * frcp((rhoa + rhob) * den)) * Gbar;
          }
          if (VISC == VISC_MON) {
            const float viss = r != 0.0f ? vdr * frcp(r) : 0.0f;
            if (!LIQDEM) vsig = cs2 + 2.0f * fabsf(viss);
            fvr = mb * (viss * vsig * alphaM * frcp(rhoa + rhob)) * Gbar;
          }
          if (ARTIF) {
            float avsig = 0.0f, avisc = 0.0f;
            if (vdr < 0.0f) {
              const float viss = r != 0.0f ? vdr * frcp(r) : 0.0f;
              avsig = cs2 - 4.0f * viss;
              avisc = viss * avsig * alpha * frcp(rhoa + rhob);
            }
            const float s = mb * avisc * Gbar;
            ax += (double)(s * dx.x);
            ay += (double)(s * dx.y);
            az += (double)(s * dx.z);
            vsig = fmaxf(vsig, avsig);
            newDt = dmin2(newDt, 0.6 * hav / (double)vsig);
          } else if (VISC == VISC_MON) {
            newDt = dmin2(newDt, 0.6 * hav / (double)vsig);
          }
          minHav = dmin2(minHav, hav);
          constexpr bool alongDr = VISC == VISC_CLE || VISC == VISC_MON;
          const double f0 = (double)(alongDr ? dx.x * fvr : dv0 * fvs);
          const double f1 = (double)(alongDr ? dx.y * fvr : dv1 * fvs);
          const double f2 = (double)(alongDr ? dx.z * fvr : dv2 * fvs);
          vx += f0;
          vy += f1;
          vz += f2;
        }
        const float pp = -mb * (pdr2a * Ga + pdr2b * Gb);  // pressure force = dx * pp
        const double q0 = (double)(pp * dx.x), q1 = (double)(pp * dx.y), q2 = (double)(pp * dx.z);
        px += q0;
        py += q1;
        pz += q2;
      }
    }
    if (minHav != inf) {
      if (!ARTIF && VISC != VISC_MON) newDt = dmin2(newDt, (0.6 * minHav) / vsigC);
      newDt = dmin2(newDt, (cdiff * (minHav * minHav)) / P.viscosity);
    }
    double fx = px + vx + ax, fy = py + vy + ay, fz = pz + vz + az;
    if (OMG) {
      const double omega = 1 + ha * om / (da.x * 3);
      S.dpo[a].z = omega;
      const double rw = 1.0 / omega;
      fx *= rw;
      fy *= rw;
      fz *= rw;
      px *= rw;
      py *= rw;
      pz *= rw;
      vx *= rw;
      vy *= rw;
      vz *= rw;
    }
    double4 f = S.frc[a], fpa = S.fp[a], fva = S.fv[a];
    f.x += fx;
    f.y += fy;
    f.z += fz;
    fpa.x += px;
    fpa.y += py;
    fpa.z += pz;
    fva.x += vx;
    fva.y += vy;
    fva.z += vz;
    S.frc[a] = f;
    S.fp[a] = fpa;
    S.fv[a] = fva;
  }
  atomicMinPos(&g->newDt_bits, newDt);
}

}  // namespace sphdem
