// FP32-compensated (FP32C) form of the MCV step.
//
// Storage: the geopotential is kept as an FP32 deviation phi' = phi - phi_ref
// from an FP64 reference geopotential (the solve's initial phi after one
// boundary pass, so that it is continuous across panel edges and its halo is
// the interpolation of its neighbours); u and v are kept in FP32.  The RK
// base state Y_n carries an FP32 compensation word per field (Y = hi + lo),
// so the long-time accumulation is done in ~48-bit precision.
//
// Arithmetic: the stencil (one-sided fits, Riemann fluxes, interior
// derivatives) runs in FP32 on E' = phi' + K, on the mass flux and on the
// winds.  The large balanced part of the pressure gradient, the MCV
// derivative of E_ref = phi_ref + phi_s, is computed once per solve in FP64.
// The reference mass term phi_ref*sqrtG stays FP64: its FP32 rounding is a
// fixed error in the mass flux, the same at every step, which the scheme's
// dissipation does not remove (MCV3 6x32^2 jet, 10 days: 2.7e-6 in h with
// it rounded to FP32, 1.1e-7 without).
// The tendencies are assembled and the RK update is done in FP64 (on gfx950
// plain FP32 and FP64 vector instructions issue at the same rate), then
// rounded back to FP32 hi/lo.  Metric terms, ghost weights and the
// mass/energy integrals stay FP64.
#pragma once

#include "mcv_device.hpp"
#include "mcv_stage_fp64.hpp"

namespace mcv {

// Static data of an FP32C solve (per resolution and reference state).
struct MeshF32 {
  int dof, n, npv, nh;
  double rdx, r2dx, rdxx;       // 1/dx, 1/(2 dx), 1/dx^2
  double dxf;                   // dx
  long long metric_pstride;
  const double *sg, *rsg, *ig11, *ig12, *ig22;  // FP64 metric (same arrays as the FP64 path)
  const D2 *mpack, *spack;      // the same terms packed per point (MeshDev::mpack, spack)
  const double *pgref;          // phi_ref * sqrtG (FP64: a rounded copy is a fixed mass-flux error)
  // {dE/dx, dE/dy, f, 0}: RN32 of the FP64 MCV derivatives (x, y) of
  // phi_ref + phi_s and of the Coriolis parameter, one 16-byte load
  const F4 *fdE;
};

struct StageArgs32 {
  // each level, compensation and accumulator as one {phi', u, v} record per point
  const F3 *in;                            // stage state
  const F3 *b;                             // base Y_n (hi)
  const F3 *c;                             // base Y_n compensation (lo)
  F3 *acc;                                 // RK4 accumulator
  F3 *o;                                   // stage output (raw interior)
  F3 *oc;                                  // compensation of the output (last stage only)
  F3 *band_raw;                            // raw band lines (hi), see band_point()
  F3 *band_raw_c;                          // raw band lines (lo), last stage only
  double c1;                               // stage increment coefficient (x dt)
};

// Arithmetic type of the FP32C stencil (LDS tiles, one-sided fits, Riemann
// fluxes, interior derivatives); level storage is FP32 whatever this is.
// FP32: an FP64 stencil on the same FP32 levels cuts the distance from the
// FP64 trajectory by ~15x on stable flows (MCV3 6x32^2 jet, 10 days: 1.1e-7
// -> 7.5e-9 in h), but doubles the LDS tile, lowers the resident workgroups
// of the FP32C kernels (8x4 MCV4 tile 6 -> 4, batched MCV3 body 8 -> 5) and
// costs 3.5% of the step; FP32 already keeps h within 1e-6 there.
using FP32CStencil = float;

template <typename T>
__device__ __forceinline__ T fit_left_t(const T *pv, int dof, T r, T r2) {
  if (dof == 3) return -((pv[2] - T(4) * pv[1] + T(3) * pv[0]) * r);
  return (-T(11) * pv[0] + T(18) * pv[1] - T(9) * pv[2] + T(2) * pv[3]) * r2;
}
template <typename T>
__device__ __forceinline__ T fit_right_t(const T *pv, int dof, T r, T r2) {
  if (dof == 3) return (pv[0] - T(4) * pv[1] + T(3) * pv[2]) * r;
  return (T(11) * pv[3] - T(18) * pv[2] + T(9) * pv[1] - T(2) * pv[0]) * r2;
}
template <typename T>
__device__ __forceinline__ T riemann_t(T fxL, T fxR, T qxL, T qxR, T lam) {
  return T(0.5) * ((fxL + fxR) - lam * (qxR - qxL));
}
template <int DOF, typename T>
__device__ __forceinline__ T cell_interior_t(const T *f, T dL, T dR, int pos, T dxf, T rdx,
                                             T rdxx) {
  if constexpr (DOF == 3) {
    return T(1.5) * ((f[2] - f[0]) * rdx) - T(0.25) * (dL + dR);
  } else {
    const T via = (f[3] - f[0]) * rdx;
    const T fxx = T(4.5) * (f[0] - f[1] - f[2] + f[3]) * rdxx;
    const bool first = pos == 1;
    const T cl = first ? T(4) : T(5), cr = first ? T(5) : T(4);
    const T a = (T(4) / T(3)) * via - (cl * dL + cr * dR) * (T(1) / T(27));
    const T c = (T(4) / T(27)) * dxf * fxx;
    return first ? a - c : a + c;
  }
}

// One (panel p, tile tx, ty) of the FP32C operator + RK stage `kind`; LDS of
// at least TileLayout<FP32CStencil, DOF, CX, CY>::kBytes.
template <int DOF, int CX, int CY>
__device__ __forceinline__ void stage_tile_f32(const MeshF32 &m, const StageArgs32 &a,
                                               const int kind, const int p, const int tx,
                                               const int ty, char *smem) {
  using C = FP32CStencil;
  using TL = TileLayout<C, DOF, CX, CY>;
  constexpr int H = DOF - 1;
  constexpr int NX = TL::NX, NY = TL::NY, TX = TL::TX, TY = TL::TY;
  const TL lds(smem);
  C *const sE = lds.sE, *const sU = lds.sU, *const sV = lds.sV, *const sPG = lds.sPG,
          *const sCU = lds.sCU, *const sCV = lds.sCV, *const lamx = lds.lamx,
          *const lamy = lds.lamy, *const rxF = lds.rxF, *const rxD = lds.rxD,
          *const rxV = lds.rxV, *const ryF = lds.ryF, *const ryD = lds.ryD, *const ryU = lds.ryU;

  const int npv = m.npv, nh = m.nh, N = m.n;
  const C rdx = (C)m.rdx, r2dx = (C)m.r2dx, rdxx = (C)m.rdxx, dxf = (C)m.dxf;
  const int cx0 = tx * CX, cy0 = ty * CY;
  const int ncx = min(CX, N - cx0), ncy = min(CY, N - cy0);
  const int i0 = cx0 * H + 1, j0 = cy0 * H + 1;
  const int nx = ncx * H + 1, ny = ncy * H + 1;
  const int TXo = ncx * H + (cx0 + ncx == N ? 1 : 0);
  const int TYo = ncy * H + (cy0 + ncy == N ? 1 : 0);
  const int lw = nx + 2 * H, lh = ny + 2 * H;
  // 32-bit point indices (every array here is below 2^32 bytes)
  const uint32_t pbase = (uint32_t)(p - 1) * (uint32_t)(nh * nh);
  const uint32_t mbase = (uint32_t)(p - 1) * (uint32_t)m.metric_pstride;
  const D2 *const mpk = m.mpack + 2 * (size_t)mbase;  // this panel's packed metric terms
  // byte offset of point (row, col) of this panel in the 12-byte record arrays,
  // from full-rate 24-bit multiplies (row, col < nh <= 6148)
  const uint32_t pb12 = 12u * pbase, nh12 = 12u * (uint32_t)nh;
  auto rec = [&](int row, int col) { return pb12 + mul24(row, nh12) + mul24(col, 12u); };

  // loads of all a thread's points first, then the derived values
  constexpr int kP0 = (TX * TY + kStageThreads - 1) / kStageThreads;
  for (int t0 = threadIdx.x; t0 < TX * TY; t0 += kP0 * blockDim.x) {
    F3 rs[kP0];
    double rpg[kP0];
    double r11[kP0], r12[kP0], r22[kP0], rsg[kP0];
    bool live[kP0];
#pragma unroll
    for (int q = 0; q < kP0; ++q) {
      const int t = t0 + q * blockDim.x;
      const int ly = div_small<TX>(t), lx = t - ly * TX;
      const bool inx = lx >= H && lx < H + nx, iny = ly >= H && ly < H + ny;
      live[q] = t < TX * TY && lx < lw && ly < lh && (inx || iny);
      if (!live[q]) continue;
      const uint32_t off = mul24(j0 - 1 + ly, nh) + (uint32_t)(i0 - 1 + lx);
      const uint32_t g = pbase + off;
      rs[q] = ldgb(a.in, rec(j0 - 1 + ly, i0 - 1 + lx));
      rpg[q] = ldg(m.pgref, g);
      const D2 ga = ldg(mpk, 2 * off), gb = ldg(mpk, 2 * off + 1);
      rsg[q] = ga.a;
      r11[q] = ga.b;
      r12[q] = gb.a;
      r22[q] = gb.b;
    }
#pragma unroll
    for (int q = 0; q < kP0; ++q) {
      if (!live[q]) continue;
      const int li = t0 + q * blockDim.x;  // = ly * TX + lx
      const int ly = div_small<TX>(li), lx = li - ly * TX;
      const float dp = rs[q].a, u = rs[q].b, v = rs[q].c;
      // metric products in FP64 (the FP32 inputs are exact in FP64), rounded once
      const double ud = u, vd = v;
      const double g11 = r11[q], g12 = r12[q], g22 = r22[q];
      const double cud = g11 * ud + g12 * vd, cvd = g12 * ud + g22 * vd;
      const double Kd = 0.5 * (cud * ud + cvd * vd);
      const C pg = (C)(rpg[q] + (double)dp * rsg[q]);
      const C cu = (C)cud, cv = (C)cvd;
      sE[li] = (C)((double)dp + Kd);  // E - E_ref
      sU[li] = u;
      sV[li] = v;
      sPG[li] = pg;
      const bool inx = lx >= H && lx < H + nx, iny = ly >= H && ly < H + ny;
      if (iny) sCU[TL::cu_at(lx, ly)] = cu;
      if (inx) sCV[TL::cv_at(lx, ly)] = cv;
      if (inx && iny) {
        const int ox = lx - H, oy = ly - H;
        lds.sSG[TL::core_at(lx, ly)] = rsg[q];
        const C phi = (C)((double)pg / rsg[q]);  // for the wave speeds only
        const int cx = div_small<H>(ox), cy = div_small<H>(oy);
        if (ox == cx * H) {
          const C sq = sqrt((C)g11 * phi);
          lamx[cx * NY + oy] = fmax(fabs(cu - sq), fabs(cu + sq));
        }
        if (oy == cy * H) {
          const C sq = sqrt((C)g22 * phi);
          lamy[cy * NX + ox] = fmax(fabs(cv - sq), fabs(cv + sq));
        }
      }
    }
  }
  __syncthreads();

  constexpr int kTasksX = (CX + 1) * NY, kTasksY = (CY + 1) * NX;
  for (int t = threadIdx.x; t < kTasksX + kTasksY; t += blockDim.x) {
    const bool isx = t < kTasksX;
    const int tt = isx ? t : t - kTasksX;
    int mb, o;
    if (isx) { mb = div_small<NY>(tt); o = tt - mb * NY; }
    else { mb = div_small<NX>(tt); o = tt - mb * NX; }
    if (isx ? (mb > ncx || o >= ny) : (mb > ncy || o >= nx)) continue;
    const int st = isx ? 1 : TX;
    const int b = isx ? (o + H) * TX + (H + mb * H) : (H + mb * H) * TX + (o + H);
    const C *sC = isx ? sCU : sCV;
    const int cst = isx ? 1 : NX;  // the same line in the contravariant wind's slots
    const int cb = isx ? TL::cu_at(H + mb * H, o + H) : TL::cv_at(o + H, H + mb * H);
    const C *sQ = isx ? sU : sV;
    const C *sW = isx ? sV : sU;
    const C lam = isx ? lamx[tt] : lamy[tt];
    const C c = sC[cb];
    const C up = c == C(0) ? C(1) : copysign(C(1), c);
    // the three flux pairs one after the other (only one pair's windows live)
    C fF, fD, fW;
    {
      C eL[DOF], eR[DOF], qL[DOF], qR[DOF];
#pragma unroll
      for (int k = 0; k < DOF; ++k) {
        const int l = b - H * st + k * st, r = b + k * st;
        eL[k] = sE[l];  eR[k] = sE[r];
        qL[k] = sQ[l];  qR[k] = sQ[r];
      }
      fF = riemann_t(fit_right_t(eL, DOF, rdx, r2dx), fit_left_t(eR, DOF, rdx, r2dx),
                     fit_right_t(qL, DOF, rdx, r2dx), fit_left_t(qR, DOF, rdx, r2dx), lam);
    }
    {
      C dL[DOF], dR[DOF], pL[DOF], pR[DOF];
#pragma unroll
      for (int k = 0; k < DOF; ++k) {
        const int l = b - H * st + k * st, r = b + k * st;
        pL[k] = sPG[l]; pR[k] = sPG[r];
        dL[k] = sPG[l] * sC[cb - H * cst + k * cst];
        dR[k] = sPG[r] * sC[cb + k * cst];
      }
      fD = riemann_t(fit_right_t(dL, DOF, rdx, r2dx), fit_left_t(dR, DOF, rdx, r2dx),
                     fit_right_t(pL, DOF, rdx, r2dx), fit_left_t(pR, DOF, rdx, r2dx), lam);
    }
    {
      C wL[DOF], wR[DOF];
#pragma unroll
      for (int k = 0; k < DOF; ++k) {
        wL[k] = sW[b - H * st + k * st];
        wR[k] = sW[b + k * st];
      }
      const C wlr = fit_right_t(wL, DOF, rdx, r2dx), wrl = fit_left_t(wR, DOF, rdx, r2dx);
      fW = riemann_t(wlr, wrl, wlr, wrl, up);
    }
    if (isx) {
      rxF[tt] = fF; rxD[tt] = fD; rxV[tt] = fW;
    } else {
      ryF[tt] = fF; ryD[tt] = fD; ryU[tt] = fW;
    }
  }
  __syncthreads();

  // global loads of all a thread's points first, then their LDS work
  constexpr int kCells = CX * CY, kSlots = H * H * kCells, kTail = NY + NX - 1;
  constexpr int kP2 = (kSlots + kTail + kStageThreads - 1) / kStageThreads;
  const bool first = kind == kRK4S1 || kind == kRK3S1;
  const bool need_in = first || kind == kRK3S2 || kind == kRK3S3;
  const bool need_acc = kind == kRK4S2 || kind == kRK4S3 || kind == kRK4S4;
  // Only the step's last stage needs Y_n to full (hi + lo) precision: the
  // intermediate stage states are rounded to FP32 anyway.
  const bool last = kind == kRK4S4 || kind == kRK3S3;
  for (int t0 = threadIdx.x; t0 < kSlots + kTail; t0 += kP2 * blockDim.x) {
    int rox[kP2], roy[kP2];
    bool live[kP2];
    float rf[kP2], rdx_[kP2], rdy_[kP2], rin[kP2], rbp[kP2], rbu[kP2], rbv[kP2], rcp[kP2],
        rcu[kP2], rcv[kP2], rap[kP2], rau[kP2], rav[kP2];
#pragma unroll
    for (int q = 0; q < kP2; ++q) {
      const int t = t0 + q * blockDim.x;
      int ox = 0, oy = 0;
      bool ok = false;
      if (t < kSlots) {
        const int cls = div_small<kCells>(t), cell = t - cls * kCells;
        const int cyl = div_small<CX>(cell), cxl = cell - cyl * CX;
        const int cly = div_small<H>(cls);
        ok = cxl < ncx && cyl < ncy;
        ox = cxl * H + (cls - cly * H);
        oy = cyl * H + cly;
      } else if (t < kSlots + NY) {
        oy = t - kSlots;
        ox = ncx * H;
        ok = ox < TXo && oy < TYo;
      } else if (t < kSlots + kTail) {
        ox = t - kSlots - NY;
        oy = ncy * H;
        ok = ox < ncx * H && oy < TYo;
      }
      live[q] = ok;
      rox[q] = ox;
      roy[q] = oy;
      if (!ok) continue;
      const uint32_t off = mul24(j0 + oy + H - 1, nh) + (uint32_t)(i0 + ox + H - 1);
      const uint32_t g = pbase + off;
      const uint32_t gb = rec(j0 + oy + H - 1, i0 + ox + H - 1);
      const F4 fde = ldg(m.fdE, g);
      rdx_[q] = fde.a;
      rdy_[q] = fde.b;
      rf[q] = fde.c;
      if (kind == kTend) continue;
      if (last) {
        const F3 cc = ldgb(a.c, gb);
        rcp[q] = cc.a;
        rcu[q] = cc.b;
        rcv[q] = cc.c;
      }
      // phi' of the stage input alone (its u, v are staged)
      if (need_in) rin[q] = ldgb(reinterpret_cast<const float *>(a.in), gb);
      if (!first) {
        const F3 bb = ldgb(a.b, gb);
        rbp[q] = bb.a;
        rbu[q] = bb.b;
        rbv[q] = bb.c;
      }
      if (need_acc) {
        const F3 aa = ldgb(a.acc, gb);
        rap[q] = aa.a;
        rau[q] = aa.b;
        rav[q] = aa.c;
      }
    }
#pragma unroll
    for (int q = 0; q < kP2; ++q) {
      if (!live[q]) continue;
      const int ox = rox[q], oy = roy[q];
      const int lx = ox + H, ly = oy + H, li = ly * TX + lx;
      C fx, dvx, vx, fy, dvy, uy;
      {
        const int mb = div_small<H>(ox), pos = ox - mb * H;
        const int r0 = mb * NY + oy;
        if (pos == 0) {
          fx = rxF[r0]; dvx = rxD[r0]; vx = rxV[r0];
        } else {
          const int b0 = ly * TX + H + mb * H;
          C fe[DOF], fd[DOF], fv[DOF];
#pragma unroll
          for (int k = 0; k < DOF; ++k) {
            fe[k] = sE[b0 + k]; fd[k] = sPG[b0 + k] * sCU[b0 + k - H * TX]; fv[k] = sV[b0 + k];
          }
          const int r1 = r0 + NY;
          fx = cell_interior_t<DOF>(fe, rxF[r0], rxF[r1], pos, dxf, rdx, rdxx);
          dvx = cell_interior_t<DOF>(fd, rxD[r0], rxD[r1], pos, dxf, rdx, rdxx);
          vx = cell_interior_t<DOF>(fv, rxV[r0], rxV[r1], pos, dxf, rdx, rdxx);
        }
      }
      {
        const int mb = div_small<H>(oy), pos = oy - mb * H;
        const int r0 = mb * NX + ox;
        if (pos == 0) {
          fy = ryF[r0]; dvy = ryD[r0]; uy = ryU[r0];
        } else {
          const int b0 = (H + mb * H) * TX + lx;
          C fe[DOF], fd[DOF], fu[DOF];
#pragma unroll
          for (int k = 0; k < DOF; ++k) {
            const int l = b0 + k * TX;
            fe[k] = sE[l]; fd[k] = sPG[l] * sCV[(H + mb * H + k) * NX + ox]; fu[k] = sU[l];
          }
          const int r1 = r0 + NX;
          fy = cell_interior_t<DOF>(fe, ryF[r0], ryF[r1], pos, dxf, rdx, rdxx);
          dvy = cell_interior_t<DOF>(fd, ryD[r0], ryD[r1], pos, dxf, rdx, rdxx);
          uy = cell_interior_t<DOF>(fu, ryU[r0], ryU[r1], pos, dxf, rdx, rdxx);
        }
      }
      const int i = i0 + ox, j = j0 + oy;
      const uint32_t gb = rec(j + H - 1, i + H - 1);
      // tendencies assembled in FP64 (spatial_operators_mod.F90:127-131)
      const double sg = lds.sSG[TL::core_at(lx, ly)], rsgv = 1.0 / sg;
      const double fc = (double)rf[q];
      const double cu = sCU[TL::cu_at(lx, ly)], cv = sCV[TL::cv_at(lx, ly)];
      const double vort = ((double)vx - (double)uy) * rsgv;
      const double kpg = -(double)dvx - (double)dvy;
      const double ku = -((double)rdx_[q] + (double)fx) + sg * (vort + fc) * cv;
      const double kv = -((double)rdy_[q] + (double)fy) - sg * (vort + fc) * cu;
      const double kph = kpg * rsgv;  // d(phi')/dt = d(phiG)/dt / sqrtG
      if (kind == kTend) {
        const uint32_t ti = ((uint32_t)(p - 1) * npv + (j - 1)) * npv + (i - 1);
        stg(a.o, ti, F3{(float)kpg, (float)ku, (float)kv});
      } else {
        // base state Y_n: hi (+ lo in the last stage) in FP64
        double bph, bu, bv;
        if (kind == kRK4S1 || kind == kRK3S1) {
          bph = (double)rin[q];
          bu = (double)sU[li];
          bv = (double)sV[li];
        } else {
          bph = (double)rbp[q];
          bu = (double)rbu[q];
          bv = (double)rbv[q];
        }
        if (last) {
          bph += (double)rcp[q];
          bu += (double)rcu[q];
          bv += (double)rcv[q];
        }
        double ph, u, v;
        if (kind == kRK4S1) {
          ph = bph + a.c1 * kph; u = bu + a.c1 * ku; v = bv + a.c1 * kv;
          stgb(a.acc, gb, F3{(float)kph, (float)ku, (float)kv});
        } else if (kind == kRK4S2 || kind == kRK4S3) {
          ph = bph + a.c1 * kph; u = bu + a.c1 * ku; v = bv + a.c1 * kv;
          stgb(a.acc, gb,
               F3{(float)((double)rap[q] + 2.0 * kph), (float)((double)rau[q] + 2.0 * ku),
                 (float)((double)rav[q] + 2.0 * kv)});
        } else if (kind == kRK4S4) {
          ph = bph + a.c1 * (((double)rap[q] + kph) * (1.0 / 6.0));
          u = bu + a.c1 * (((double)rau[q] + ku) * (1.0 / 6.0));
          v = bv + a.c1 * (((double)rav[q] + kv) * (1.0 / 6.0));
        } else if (kind == kRK3S1) {
          ph = bph + a.c1 * kph; u = bu + a.c1 * ku; v = bv + a.c1 * kv;
        } else if (kind == kRK3S2) {  // Yn + 1/4 (Y1 - Yn) + 1/4 dt k
          ph = bph + 0.25 * ((double)rin[q] - bph) + a.c1 * kph;
          u = bu + 0.25 * ((double)sU[li] - bu) + a.c1 * ku;
          v = bv + 0.25 * ((double)sV[li] - bv) + a.c1 * kv;
        } else {  // kRK3S3: Yn + 2/3 (Y2 - Yn) + 2/3 dt k
          ph = bph + (2.0 / 3.0) * ((double)rin[q] - bph) + a.c1 * kph;
          u = bu + (2.0 / 3.0) * ((double)sU[li] - bu) + a.c1 * ku;
          v = bv + (2.0 / 3.0) * ((double)sV[li] - bv) + a.c1 * kv;
        }
        const float hph = (float)ph, hu = (float)u, hv = (float)v;
        const F3 hi{hph, hu, hv};
        // the rounding residue (Kahan word), kept by the last stage only
        const F3 lo{(float)(ph - (double)hph), (float)(u - (double)hu), (float)(v - (double)hv)};
        stgb(a.o, gb, hi);
        if (a.oc) stgb(a.oc, gb, lo);
        // raw band records (and their residues in the last stage) for the boundary pass
        if (i <= H + 1 || i >= npv - H || j <= H + 1 || j >= npv - H) {
          // one record per side the point is near (a corner point is near two)
          auto put = [&](uint32_t bp) {
            stg(a.band_raw, bp, hi);
            if (a.oc) stg(a.band_raw_c, bp, lo);
          };
          if (i <= H + 1) put(band_point(npv, H, p, 0, i, j));
          if (i >= npv - H) put(band_point(npv, H, p, 1, npv + 1 - i, j));
          if (j <= H + 1) put(band_point(npv, H, p, 2, j, i));
          if (j >= npv - H) put(band_point(npv, H, p, 3, npv + 1 - j, i));
        }
      }
    }
  }
}

template <int DOF, int KIND, int CX, int CY>
__global__ __launch_bounds__(kStageThreads) void stage_kernel_f32(MeshF32 m, StageArgs32 a,
                                                                  int tiles_x) {
  __shared__ __align__(16) char smem[TileLayout<FP32CStencil, DOF, CX, CY>::kBytes];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = L / 6;
  stage_tile_f32<DOF, CX, CY>(m, a, KIND, L % 6 + 1, tile % tiles_x, tile / tiles_x, smem);
}

// correct_bdy_ghost on an FP32C level: same algorithm as the FP64 boundary
// kernel, values in FP32 storage, arithmetic and ghost weights in FP64.  The
// deviation phi' is interpolated and averaged like phi (both operations are
// linear and phi_ref is already continuous across edges).  For the base state
// (cpu != NULL) values are hi + lo and every written point gets its residue.
struct BoundaryArgs32 {
  F3 *lev;                 // level to finish: {phi', u, v} records
  F3 *comp;                // its compensation words (base state only, else NULL)
  const F3 *band_raw, *band_raw_c;  // raw band records; residues for the base state
};

template <int DOF>
__device__ __forceinline__ void boundary_item_f32(const MeshDev &m, const BoundaryArgs32 &ba,
                                                  const int t) {
  constexpr int H = DOF - 1;
  F3 *const lev = ba.lev, *const comp = ba.comp;
  const int npv = m.npv, nh = m.nh;
  const double dx = m.dx;
  const int nghost = 6 * 4 * H * npv, nedge = 12 * npv;
  // raw value of band point (panel q, side, depth d, along index al): hi, plus
  // lo for the base state
  auto raw = [&](int q, int side, int d, int al, double &ph, double &uu, double &vv) {
    const uint32_t bp = band_point(npv, H, q, side, d, al);
    const F3 r = ldg(ba.band_raw, bp);
    ph = r.a; uu = r.b; vv = r.c;
    if (comp) {
      const F3 c = ldg(ba.band_raw_c, bp);
      ph += (double)c.a;
      uu += (double)c.b;
      vv += (double)c.c;
    }
  };
  auto store = [&](long long g, double ph, double uu, double vv) {
    const float a0 = (float)ph, a1 = (float)uu, a2 = (float)vv;
    lev[g] = F3{a0, a1, a2};
    if (comp) comp[g] = F3{(float)(ph - (double)a0), (float)(uu - (double)a1), (float)(vv - (double)a2)};
  };
  if (t < nghost) {  // one ghost point: (panel, side, depth k, s), s fastest
    const int s = t % npv + 1, r = t / npv, k = r % H + 1, ps = r / H;
    const int side = ps % 4, p = ps / 4 + 1;
    const HaloSrc hs = kHalo[p - 1][side];
    const int src = hs.col ? (hs.far ? 1 : 0) : (hs.far ? 3 : 2);
    const int c = m.giref[(k - 1) * npv + (s - 1)];
    const double cf = m.gcoef[(k - 1) * npv + (s - 1)];
    double qp[DOF], qz[DOF], qm[DOF];
    for (int l = 0; l < DOF; ++l) {
      const int tpos = (c - 1) * H + 1 + l;
      const int al = hs.rev ? npv + 1 - tpos : tpos;
      double ph, uu, vv;
      raw(hs.q, src, k + 1, al, ph, uu, vv);
      qp[l] = ph;
      apply2(m.p2sp_band + 4 * band_point(npv, H, hs.q, src, k + 1, al), uu, vv, qz[l], qm[l]);
    }
    int hi, hj;
    if (side == 0) { hi = 1 - k; hj = s; }
    else if (side == 1) { hi = npv + k; hj = s; }
    else if (side == 2) { hi = s; hj = 1 - k; }
    else { hi = s; hj = npv + k; }
    const long long g = pidx(nh, H, p, hi, hj);
    const double zh = ghost_interp<DOF>(qz, cf, dx), mh = ghost_interp<DOF>(qm, cf, dx);
    double uh, vh;
    apply2(m.sp2p_ghost + 4 * ghost_point(npv, H, p, side, k, s), zh, mh, uh, vh);
    store(g, ghost_interp<DOF>(qp, cf, dx), uh, vh);
  } else if (t < nghost + nedge) {  // one point of one cube edge
    const int s = (t - nghost) % npv + 1, e = (t - nghost) / npv;
    if (s > 1 && s < npv) {
      const EdgePair ep = kEdges[e];
      const int sb = ep.rev ? npv + 1 - s : s;
      int ia, ja, ib, jb;
      edge_ij(npv, ep.sa, s, ia, ja);
      edge_ij(npv, ep.sb, sb, ib, jb);
      const long long ga = pidx(nh, H, ep.pa, ia, ja), gb = pidx(nh, H, ep.pb, ib, jb);
      double pa, ua, va, pb, ub, vb;
      raw(ep.pa, ep.sa, 1, s, pa, ua, va);
      raw(ep.pb, ep.sb, 1, sb, pb, ub, vb);
      double za, ma, zb, mb;
      apply2(m.p2sp_band + 4 * band_point(npv, H, ep.pa, ep.sa, 1, s), ua, va, za, ma);
      apply2(m.p2sp_band + 4 * band_point(npv, H, ep.pb, ep.sb, 1, sb), ub, vb, zb, mb);
      const double ph = 0.5 * (pa + pb), zu = 0.5 * (za + zb), mu = 0.5 * (ma + mb);
      double u1, v1, u2, v2;
      apply2(m.sp2p_edge + 4 * edge_point(npv, ep.pa, ep.sa, s), zu, mu, u1, v1);
      apply2(m.sp2p_edge + 4 * edge_point(npv, ep.pb, ep.sb, sb), zu, mu, u2, v2);
      store(ga, ph, u1, v1);
      store(gb, ph, u2, v2);
    }
  } else {
    const int vx = t - nghost - nedge;
    double ph[3], z[3], mr[3];
    long long gg[3];
    uint32_t eg[3];
    for (int l = 0; l < 3; ++l) {
      const VtxPt vp = kVerts[vx][l];
      const int i = vp.ci ? npv : 1, j = vp.cj ? npv : 1;
      const int side = vp.ci ? 1 : 0;
      double uu, vv;
      raw(vp.p, side, 1, j, ph[l], uu, vv);
      gg[l] = pidx(nh, H, vp.p, i, j);
      eg[l] = edge_point(npv, vp.p, side, j);
      apply2(m.p2sp_band + 4 * band_point(npv, H, vp.p, side, 1, j), uu, vv, z[l], mr[l]);
    }
    const double pv = (ph[0] + ph[1] + ph[2]) / 3.0;
    const double zv = (z[0] + z[1] + z[2]) / 3.0;
    const double mv = (mr[0] + mr[1] + mr[2]) / 3.0;
    for (int l = 0; l < 3; ++l) {
      double uu, vv;
      apply2(m.sp2p_edge + 4 * eg[l], zv, mv, uu, vv);
      store(gg[l], pv, uu, vv);
    }
  }
}

template <int DOF>
__global__ __launch_bounds__(256) void boundary_kernel_f32(MeshDev m, BoundaryArgs32 a) {
  const int nitems = boundary_items(DOF, m.npv);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nitems; t += gridDim.x * blockDim.x)
    boundary_item_f32<DOF>(m, a, t);
}

// pgref = phi_ref * sqrtG on every point.
static __global__ void pgref_kernel(MeshDev m, const double *phi_ref, double *pgref) {
  const long long per = (long long)m.nh * m.nh;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < 6 * per;
       g += (long long)gridDim.x * blockDim.x) {
    const long long p = g / per;
    pgref[g] = phi_ref[g] * m.sqrtG[p * m.metric_pstride + (g - p * per)];
  }
}

// FP64 band copy of a finished level (to run the boundary pass on it).
static __global__ void band_copy_kernel(MeshDev m, const D2 *pu, const double *v,
                                        double *band) {
  const int npv = m.npv, H = m.dof - 1;
  const int n = 6 * 4 * (H + 1) * npv;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int al = t % npv + 1, r = t / npv, d = r % (H + 1) + 1, ps = r / (H + 1);
    const int side = ps % 4, p = ps / 4 + 1;
    int i, j;
    band_ij(npv, side, d, al, i, j);
    const long long g = pidx(m.nh, H, p, i, j);
    const D2 x = pu[g];
    band[band_slot(npv, H, p, side, d, 0, al)] = x.a;
    band[band_slot(npv, H, p, side, d, 1, al)] = x.b;
    band[band_slot(npv, H, p, side, d, 2, al)] = v[g];
  }
}

// FP64 MCV derivatives (x and y) of E_ref = phi_ref + phi_s with the
// dissipation term left out: the part of dE/dx, dE/dy that does not depend on
// the evolving state, rounded to FP32 beside the rounded Coriolis parameter.
// One thread per owned point; runs once per solve.
template <int DOF>
__global__ void ref_gradient_kernel(MeshDev m, const double *phi_ref, F4 *fdE) {
  constexpr int H = DOF - 1;
  const int npv = m.npv, nh = m.nh;
  const long long per = (long long)nh * nh;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < 6LL * npv * npv;
       t += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(t / ((long long)npv * npv)) + 1;
    const int rem = (int)(t - (long long)(p - 1) * npv * npv);
    const int j = rem / npv + 1, i = rem - (j - 1) * npv + 1;
    auto E = [&](int ii, int jj) {
      const long long g = pidx(nh, H, p, ii, jj);
      return phi_ref[g] + m.phi_s[(long long)(p - 1) * m.phis_pstride + (g - (p - 1) * per)];
    };
    // derivative along one direction of the line through (i, j)
    auto deriv = [&](bool xdir) {
      const int c = xdir ? i : j;
      const int pos = (c - 1) % H;
      const int bL = c - pos;  // left cell end on the line
      auto val = [&](int q) { return xdir ? E(q, j) : E(i, q); };
      auto central = [&](int b) {  // 0.5*(polyfitR(left cell) + polyfitL(right cell)) at b
        double l[DOF], r[DOF];
        for (int k = 0; k < DOF; ++k) { l[k] = val(b - H + k); r[k] = val(b + k); }
        return 0.5 * (fit_right<DOF>(l, m.dv) + fit_left<DOF>(r, m.dv));
      };
      const double dL = central(bL);
      if (pos == 0) return dL;
      double f[DOF];
      for (int k = 0; k < DOF; ++k) f[k] = val(bL + k);
      return cell_interior<DOF>(f, dL, central(bL + H), pos, m.dx, m.dv);
    };
    const long long g = pidx(nh, H, p, i, j);
    fdE[g] = F4{(float)deriv(true), (float)deriv(false), (float)m.f[g], 0.0f};
  }
}

// Splits the FP64 initial state into the FP32C state: phi' = phi - phi_ref
// (zero in the interior, the halo-and-edge correction elsewhere), u and v,
// each rounded to FP32 with its residue as the compensation word.
static __global__ void split_state_kernel(long long np, const double *phi, const double *phi_ref,
                                   const double *u, const double *v, F3 *hi, F3 *lo) {
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < np;
       g += (long long)gridDim.x * blockDim.x) {
    const double d = phi[g] - phi_ref[g];
    const float hd = (float)d, hu = (float)u[g], hvv = (float)v[g];
    hi[g] = F3{hd, hu, hvv};
    lo[g] = F3{(float)(d - (double)hd), (float)(u[g] - (double)hu), (float)(v[g] - (double)hvv)};
  }
}

// Reconstructs the FP64 state of an FP32C level: phi = phi_ref + hi + lo,
// u, v = hi + lo.
static __global__ void join_state_kernel(long long np, const double *phi_ref, const F3 *hi,
                                         const F3 *lo, double *phi, double *u, double *v) {
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < np;
       g += (long long)gridDim.x * blockDim.x) {
    const F3 h = hi[g], c = lo[g];
    phi[g] = phi_ref[g] + ((double)h.a + (double)c.a);
    u[g] = (double)h.b + (double)c.b;
    v[g] = (double)h.c + (double)c.c;
  }
}

}  // namespace mcv
