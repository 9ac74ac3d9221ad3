// FP64 stage kernel: spatial operator + RK stage update on one (panel, tile).
//
// Phase 0 stages the cross-shaped tile plus its one-cell halo of (phi, u, v)
// in LDS together with the derived E = phi + phi_s + K, phiG = phi*sqrtG and
// the contravariant winds, and the wave speeds at cell ends.  Phase 1 computes
// the Riemann derivative at every cell end of the tile once (the reference
// computes each interface twice, from both cells).  Phase 2 forms the
// interior-point derivatives, the vorticity/Coriolis tendencies and the RK
// update, and writes the raw updated interior plus a raw copy of the H+1
// lines nearest each panel side (the band) for the boundary pass.  Every loop runs over compile-time
// extents, so no index needs a run-time integer division.
#pragma once

#include "mcv_device.hpp"

namespace mcv {

constexpr int kStageThreads = 256;

// LDS layout of one FP64 tile of CX x CY cells, carved from a byte buffer so
// that one kernel can host tiles of several orders and precisions.
template <typename T, int DOF, int CX, int CY>
struct TileLayout {
  static constexpr int H = DOF - 1;
  static constexpr int NX = CX * H + 1;  // core extents incl. the far cell end
  static constexpr int NY = CY * H + 1;
  static constexpr int TX = NX + 2 * H;  // staged extents incl. the halo cell
  static constexpr int TY = NY + 2 * H;
  static constexpr int kPts = TX * TY;
  static constexpr int kEx = (CX + 1) * NY;  // x-sweep cell ends: (CX+1) per core row
  static constexpr int kEy = (CY + 1) * NX;  // y-sweep: (CY+1) per core column
  // The contravariant winds are read only along their own sweep: contraU on
  // the core rows (NY x TX), contraV on the core columns (TY x NX).  The
  // cell-end wave speeds share the storage of the first flux of their sweep:
  // each phase-1 task reads its wave speed before it writes that flux.
  static constexpr int kCU = NY * TX, kCV = TY * NX;
  // sqrtG at the core points (FP64 whatever T is), staged for phase 2
  static constexpr int kCore = NX * NY;
  static constexpr size_t kBytes =
      sizeof(double) * kCore + sizeof(T) * (4 * kPts + kCU + kCV + 3 * kEx + 3 * kEy);
  double *sSG;
  T *sE, *sU, *sV, *sPG, *sCU, *sCV, *lamx, *lamy, *rxF, *rxD, *rxV, *ryF, *ryD, *ryU;
  __device__ __forceinline__ explicit TileLayout(char *smem) {
    sSG = reinterpret_cast<double *>(smem);
    T *q = reinterpret_cast<T *>(smem + sizeof(double) * kCore);
    sE = q; q += kPts; sU = q; q += kPts; sV = q; q += kPts; sPG = q; q += kPts;
    sCU = q; q += kCU; sCV = q; q += kCV;
    rxF = q; q += kEx; rxD = q; q += kEx; rxV = q; q += kEx;
    ryF = q; q += kEy; ryD = q; q += kEy; ryU = q;
    lamx = rxF; lamy = ryF;
  }
  // contraU / contraV slots of staged point (lx, ly): a core row / column
  static __device__ __forceinline__ int cu_at(int lx, int ly) { return (ly - H) * TX + lx; }
  static __device__ __forceinline__ int cv_at(int lx, int ly) { return ly * NX + (lx - H); }
  static __device__ __forceinline__ int core_at(int lx, int ly) {
    return (ly - H) * NX + (lx - H);
  }
};

// One (panel p, tile tx, ty) of the spatial operator + RK stage `kind`.  The
// caller passes LDS of at least TileLayout<double, DOF, CX, CY>::kBytes and
// synchronises before reusing it.
template <int DOF, int CX, int CY>
__device__ __forceinline__ void stage_tile(const MeshDev &m, const StageArgs &a, const int kind,
                                           const int p, const int tx, const int ty, char *smem) {
  using TL = TileLayout<double, DOF, CX, CY>;
  constexpr int H = DOF - 1;
  constexpr int NX = TL::NX, NY = TL::NY, TX = TL::TX;
  constexpr int TY = TL::TY;
  const TL lds(smem);
  double *const sE = lds.sE, *const sU = lds.sU, *const sV = lds.sV, *const sPG = lds.sPG,
               *const sCU = lds.sCU, *const sCV = lds.sCV, *const lamx = lds.lamx,
               *const lamy = lds.lamy, *const rxF = lds.rxF, *const rxD = lds.rxD,
               *const rxV = lds.rxV, *const ryF = lds.ryF, *const ryD = lds.ryD,
               *const ryU = lds.ryU;

  const int npv = m.npv, nh = m.nh, N = m.n;
  const double dx = m.dx;
  const MCVDivisors dv = m.dv;
  const int cx0 = tx * CX, cy0 = ty * CY;
  const int ncx = min(CX, N - cx0), ncy = min(CY, N - cy0);
  const int i0 = cx0 * H + 1, j0 = cy0 * H + 1;
  const int nx = ncx * H + 1, ny = ncy * H + 1;           // core incl. far cell end
  const int TXo = ncx * H + (cx0 + ncx == N ? 1 : 0);     // owned points
  const int TYo = ncy * H + (cy0 + ncy == N ? 1 : 0);
  const int lw = nx + 2 * H, lh = ny + 2 * H;
  // 32-bit point indices (every array here is below 2^32 bytes)
  const uint32_t pbase = (uint32_t)(p - 1) * (uint32_t)(nh * nh);
  const uint32_t mbase = (uint32_t)(p - 1) * (uint32_t)m.metric_pstride;
  const uint32_t sbase = (uint32_t)(p - 1) * (uint32_t)m.phis_pstride;
  const D2 *const mpk = m.mpack + 2 * (size_t)mbase;  // this panel's packed metric terms

  // Phase 0: stage the cross-shaped tile + halo; derived fields per point.
  // Every thread first issues the loads of all its points, then derives and
  // stores, so that its loads are in flight together.
  constexpr int kP0 = (TX * TY + kStageThreads - 1) / kStageThreads;
  for (int t0 = threadIdx.x; t0 < TX * TY; t0 += kP0 * blockDim.x) {
    double rphi[kP0], ru[kP0], rv[kP0], rsg[kP0], r11[kP0], r12[kP0], r22[kP0], rps[kP0];
    double rcu[kP0], rcv[kP0];
    bool live[kP0];
#pragma unroll
    for (int q = 0; q < kP0; ++q) {
      const int t = t0 + q * blockDim.x;
      const int ly = t / TX, lx = t - ly * TX;
      const bool inx = lx >= H && lx < H + nx, iny = ly >= H && ly < H + ny;
      // outside the tile, or a halo corner (never read)
      live[q] = t < TX * TY && lx < lw && ly < lh && (inx || iny);
      if (!live[q]) continue;
      const uint32_t off = (uint32_t)((j0 - 1 + ly) * nh + (i0 - 1 + lx));
      const uint32_t g = pbase + off;
      const D2 pu = ldg(a.in_pu, g);
      rphi[q] = pu.a;
      ru[q] = pu.b;
      rv[q] = ldg(a.in_v, g);
      const D2 ga = ldg(mpk, 2 * off), gb = ldg(mpk, 2 * off + 1);
      rsg[q] = ga.a;
      r11[q] = ga.b;
      r12[q] = gb.a;
      r22[q] = gb.b;
      rps[q] = m.phis_zero ? 0.0 : ldg(m.phi_s, sbase + off);
      if (a.cu0) {
        rcu[q] = ldg(a.cu0, g);
        rcv[q] = ldg(a.cv0, g);
      }
    }
#pragma unroll
    for (int q = 0; q < kP0; ++q) {
      if (!live[q]) continue;
      const int li = t0 + q * blockDim.x;  // = ly * TX + lx
      const int ly = li / TX, lx = li - ly * TX;
      const double phi = rphi[q], u = ru[q], v = rv[q], g11 = r11[q], g12 = r12[q],
                   g22 = r22[q];
      double cu, cv;
      if (a.cu0) {
        cu = rcu[q];
        cv = rcv[q];
      } else {  // cov2contrav (projection_mod.F90:138-148)
        cu = g11 * u + g12 * v;
        cv = g12 * u + g22 * v;
      }
      const double K = 0.5 * (cu * u + cv * v);  // spatial_operators_mod.F90:60-61
      sE[li] = phi + rps[q] + K;
      sU[li] = u;
      sV[li] = v;
      sPG[li] = phi * rsg[q];
      const bool inx = lx >= H && lx < H + nx, iny = ly >= H && ly < H + ny;
      if (iny) sCU[TL::cu_at(lx, ly)] = cu;
      if (inx) sCV[TL::cv_at(lx, ly)] = cv;
      if (inx && iny) {
        const int ox = lx - H, oy = ly - H;
        lds.sSG[TL::core_at(lx, ly)] = rsg[q];
        if (ox % H == 0) lamx[(ox / H) * NY + oy] = wave_speed(cu, phi, g11);
        if (oy % H == 0) lamy[(oy / H) * NX + ox] = wave_speed(cv, phi, g22);
      }
    }
  }
  __syncthreads();

  // Phase 1: Riemann derivatives at every cell end of the core, x then y.
  constexpr int kTasksX = (CX + 1) * NY, kTasksY = (CY + 1) * NX;
  for (int t = threadIdx.x; t < kTasksX + kTasksY; t += blockDim.x) {
    const bool isx = t < kTasksX;
    const int tt = isx ? t : t - kTasksX;
    int mb, o;
    if (isx) { mb = tt / NY; o = tt - mb * NY; }
    else { mb = tt / NX; o = tt - mb * NX; }
    if (isx ? (mb > ncx || o >= ny) : (mb > ncy || o >= nx)) continue;
    const int st = isx ? 1 : TX;  // stride along the sweep
    const int b = isx ? (o + H) * TX + (H + mb * H) : (H + mb * H) * TX + (o + H);
    const double *sC = isx ? sCU : sCV;
    const int cst = isx ? 1 : NX;  // the same line in the contravariant wind's slots
    const int cb = isx ? TL::cu_at(H + mb * H, o + H) : TL::cv_at(o + H, H + mb * H);
    const double *sQ = isx ? sU : sV;   // q of the gradient flux
    const double *sW = isx ? sV : sU;   // upwinded cross wind
    const double lam = isx ? lamx[tt] : lamy[tt];
    const double up = upwind_sign(sC[cb]);
    // the three flux pairs one after the other, each from its own LDS reads,
    // so that only one pair's windows are live at a time
    double fF, fD, fW;
    {
      double eL[DOF], eR[DOF], qL[DOF], qR[DOF];
#pragma unroll
      for (int k = 0; k < DOF; ++k) {
        const int l = b - H * st + k * st, r = b + k * st;
        eL[k] = sE[l];  eR[k] = sE[r];
        qL[k] = sQ[l];  qR[k] = sQ[r];
      }
      fF = riemann(fit_right<DOF>(eL, dv), fit_left<DOF>(eR, dv), fit_right<DOF>(qL, dv),
                   fit_left<DOF>(qR, dv), lam);
    }
    {
      double dL[DOF], dR[DOF], pL[DOF], pR[DOF];
#pragma unroll
      for (int k = 0; k < DOF; ++k) {
        const int l = b - H * st + k * st, r = b + k * st;
        pL[k] = sPG[l]; pR[k] = sPG[r];
        dL[k] = sPG[l] * sC[cb - H * cst + k * cst];
        dR[k] = sPG[r] * sC[cb + k * cst];
      }
      fD = riemann(fit_right<DOF>(dL, dv), fit_left<DOF>(dR, dv), fit_right<DOF>(pL, dv),
                   fit_left<DOF>(pR, dv), lam);
    }
    {
      double wL[DOF], wR[DOF];
#pragma unroll
      for (int k = 0; k < DOF; ++k) {
        wL[k] = sW[b - H * st + k * st];
        wR[k] = sW[b + k * st];
      }
      const double wlr = fit_right<DOF>(wL, dv), wrl = fit_left<DOF>(wR, dv);
      fW = riemann(wlr, wrl, wlr, wrl, up);
    }
    if (isx) {
      rxF[tt] = fF; rxD[tt] = fD; rxV[tt] = fW;
    } else {
      ryF[tt] = fF; ryD[tt] = fD; ryU[tt] = fW;
    }
  }
  __syncthreads();

  // Phase 2: derivatives at owned points, tendencies, RK update.  Points are
  // visited class by class (position within the cell in x and in y), cells
  // fastest, so the lanes of a wave take the same branches and read LDS with
  // a stride of H points (no bank conflicts); the far cell-end column and row
  // owned by the last tiles follow.
  // The global loads of all a thread's points are issued before their LDS
  // work.
  constexpr int kCells = CX * CY, kSlots = H * H * kCells, kTail = NY + NX - 1;
  constexpr int kP2 = (kSlots + kTail + kStageThreads - 1) / kStageThreads;
  const bool need_base = !(kind == kTend || kind == kRK4S1 || kind == kRK3S1);
  const bool need_acc = kind == kRK4S2 || kind == kRK4S3 || kind == kRK4S4;
  for (int t0 = threadIdx.x; t0 < kSlots + kTail; t0 += kP2 * blockDim.x) {
    int rox[kP2], roy[kP2];
    bool live[kP2];
    double rf[kP2], rbp[kP2], rbu[kP2], rbv[kP2], rap[kP2], rau[kP2],
        rav[kP2];
#pragma unroll
    for (int q = 0; q < kP2; ++q) {
      const int t = t0 + q * blockDim.x;
      int ox = 0, oy = 0;
      bool ok = false;
      if (t < kSlots) {
        const int cls = t / kCells, cell = t - cls * kCells;
        const int cyl = cell / CX, cxl = cell - cyl * CX;
        ok = cxl < ncx && cyl < ncy;
        ox = cxl * H + cls % H;
        oy = cyl * H + cls / H;
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
      const uint32_t off = (uint32_t)((j0 + oy + H - 1) * nh + (i0 + ox + H - 1));
      const uint32_t g = pbase + off;
      rf[q] = ldg(m.f, g);
      if (need_base) {
        const D2 bpu = ldg(a.b_pu, g);
        rbp[q] = bpu.a;
        rbu[q] = bpu.b;
        rbv[q] = ldg(a.b_v, g);
      }
      if (need_acc) {
        const D2 pu = ldg(a.acc_pu, g);
        rap[q] = pu.a;
        rau[q] = pu.b;
        rav[q] = ldg(a.acc_v, g);
      }
    }
#pragma unroll
    for (int q = 0; q < kP2; ++q) {
      if (!live[q]) continue;
      const int ox = rox[q], oy = roy[q];
      const int lx = ox + H, ly = oy + H, li = ly * TX + lx;
      double fx, dvx, vx, fy, dvy, uy;
      {  // x sweep
        const int mb = ox / H, pos = ox - mb * H;
        const int r0 = mb * NY + oy;
        if (pos == 0) {
          fx = rxF[r0]; dvx = rxD[r0]; vx = rxV[r0];
        } else {
          const int b0 = ly * TX + H + mb * H;
          double fe[DOF], fd[DOF], fv[DOF];
#pragma unroll
          for (int k = 0; k < DOF; ++k) {
            fe[k] = sE[b0 + k]; fd[k] = sPG[b0 + k] * sCU[b0 + k - H * TX]; fv[k] = sV[b0 + k];
          }
          const int r1 = r0 + NY;
          fx = cell_interior<DOF>(fe, rxF[r0], rxF[r1], pos, dx, dv);
          dvx = cell_interior<DOF>(fd, rxD[r0], rxD[r1], pos, dx, dv);
          vx = cell_interior<DOF>(fv, rxV[r0], rxV[r1], pos, dx, dv);
        }
      }
      {  // y sweep
        const int mb = oy / H, pos = oy - mb * H;
        const int r0 = mb * NX + ox;
        if (pos == 0) {
          fy = ryF[r0]; dvy = ryD[r0]; uy = ryU[r0];
        } else {
          const int b0 = (H + mb * H) * TX + lx;
          double fe[DOF], fd[DOF], fu[DOF];
#pragma unroll
          for (int k = 0; k < DOF; ++k) {
            const int l = b0 + k * TX;
            fe[k] = sE[l]; fd[k] = sPG[l] * sCV[(H + mb * H + k) * NX + ox]; fu[k] = sU[l];
          }
          const int r1 = r0 + NX;
          fy = cell_interior<DOF>(fe, ryF[r0], ryF[r1], pos, dx, dv);
          dvy = cell_interior<DOF>(fd, ryD[r0], ryD[r1], pos, dx, dv);
          uy = cell_interior<DOF>(fu, ryU[r0], ryU[r1], pos, dx, dv);
        }
      }
      const int i = i0 + ox, j = j0 + oy;
      const uint32_t g = pbase + (uint32_t)((j + H - 1) * nh + (i + H - 1));
      const double sg = lds.sSG[TL::core_at(lx, ly)];
      const Recip sgr{sg, 1.0 / sg};  // the IEEE quotient, as the host forms RN(1/sqrtG)
      const double fc = rf[q];
      const double cu = sCU[TL::cu_at(lx, ly)], cv = sCV[TL::cv_at(lx, ly)];
      // spatial_operators_mod.F90:127-131
      const double vort = div_by(vx - uy, sgr);
      const double kpg = -dvx - dvy;
      const double ku = -fx + sg * (vort + fc) * cv;
      const double kv = -fy - sg * (vort + fc) * cu;

      if (kind == kTend) {
        const uint32_t ti = ((uint32_t)(p - 1) * npv + (j - 1)) * npv + (i - 1);
        stg(a.o_pu, ti, D2{kpg, ku});
        stg(a.o_v, ti, kv);
      } else {
        double pgn, un, vn;  // base state Y_n
        if (kind == kRK4S1 || kind == kRK3S1) {
          pgn = sPG[li]; un = sU[li]; vn = sV[li];
        } else {
          pgn = rbp[q] * sg; un = rbu[q]; vn = rbv[q];
        }
        double pg, u, v;
        if (kind == kRK4S1) {
          pg = pgn + a.c1 * kpg; u = un + a.c1 * ku; v = vn + a.c1 * kv;
          stg(a.acc_pu, g, D2{kpg, ku});
          stg(a.acc_v, g, kv);
        } else if (kind == kRK4S2 || kind == kRK4S3) {
          pg = pgn + a.c1 * kpg; u = un + a.c1 * ku; v = vn + a.c1 * kv;
          stg(a.acc_pu, g, D2{rap[q] + 2.0 * kpg, rau[q] + 2.0 * ku});
          stg(a.acc_v, g, rav[q] + 2.0 * kv);
        } else if (kind == kRK4S4) {
          constexpr Recip k6{6.0, 1.0 / 6.0};
          const double tpg = div_by(rap[q] + kpg, k6);
          const double tu = div_by(rau[q] + ku, k6);
          const double tv = div_by(rav[q] + kv, k6);
          pg = pgn + a.c1 * tpg; u = un + a.c1 * tu; v = vn + a.c1 * tv;
        } else if (kind == kRK3S1) {
          pg = pgn + a.c1 * kpg; u = un + a.c1 * ku; v = vn + a.c1 * kv;
        } else if (kind == kRK3S2) {
          pg = 0.75 * pgn + 0.25 * sPG[li] + a.c1 * kpg;
          u = 0.75 * un + 0.25 * sU[li] + a.c1 * ku;
          v = 0.75 * vn + 0.25 * sV[li] + a.c1 * kv;
        } else {  // kRK3S3
          constexpr Recip k3{3.0, 1.0 / 3.0};
          pg = div_by(pgn, k3) + 2.0 / 3.0 * sPG[li] + a.c1 * kpg;
          u = div_by(un, k3) + 2.0 / 3.0 * sU[li] + a.c1 * ku;
          v = div_by(vn, k3) + 2.0 / 3.0 * sV[li] + a.c1 * kv;
        }
        const double phi = div_by(pg, sgr);  // time_scheme_mod.F90:81
        stg(a.o_pu, g, D2{phi, u});
        stg(a.o_v, g, v);
        // raw copy of the lines nearest each side for the boundary pass
        if (i <= H + 1 || i >= npv - H || j <= H + 1 || j >= npv - H) {
          const double vals[3] = {phi, u, v};
#pragma unroll
          for (int fl = 0; fl < 3; ++fl) {
            if (i <= H + 1) stg(a.band_raw, band_slot(npv, H, p, 0, i, fl, j), vals[fl]);
            if (i >= npv - H) stg(a.band_raw, band_slot(npv, H, p, 1, npv + 1 - i, fl, j), vals[fl]);
            if (j <= H + 1) stg(a.band_raw, band_slot(npv, H, p, 2, j, fl, i), vals[fl]);
            if (j >= npv - H) stg(a.band_raw, band_slot(npv, H, p, 3, npv + 1 - j, fl, i), vals[fl]);
          }
        }
      }
    }
  }
}

// Single-solve stage kernel: one workgroup per (panel, tile), blocks of one
// XCD on neighbouring tiles.
template <int DOF, int KIND, int CX, int CY>
__global__ __launch_bounds__(kStageThreads, 6) void stage_kernel(MeshDev m, StageArgs a,
                                                              int tiles_x) {
  __shared__ __align__(16) char smem[TileLayout<double, DOF, CX, CY>::kBytes];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = L / 6;
  stage_tile<DOF, CX, CY>(m, a, KIND, L % 6 + 1, tile % tiles_x, tile / tiles_x, smem);
}

}  // namespace mcv
