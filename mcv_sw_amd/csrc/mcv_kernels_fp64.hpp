// FP64 kernels of the MCV step on gfx950.
//
// One RK stage is two launches:
//   stage_kernel     per (panel, tile): stage the tile and its one-cell halo
//                    of (phi, u, v) in LDS with the derived E, phiG, contra
//                    winds; Riemann derivatives at cell ends; interior
//                    derivatives; tendencies; RK update.  Writes the raw
//                    updated interior and a raw copy of the four edge lines.
//   boundary_kernel  per (panel, side, along-edge point): ghost interpolation
//                    into the halo and edge/vertex averaging (the reference's
//                    correct_bdy_ghost), reading the raw copy so that no
//                    thread reads a point another thread rewrites.
// phiG is never stored: every state keeps phiG = phi*sqrtG exactly
// (time_scheme_mod.F90:85), so it is recomputed from phi where needed.
#pragma once

#include "mcv_device.hpp"

namespace mcv {

template <int DOF> struct Tile;
template <> struct Tile<3> { static constexpr int C = 8; };  // 8x8 cells = 16x16 points
template <> struct Tile<4> { static constexpr int C = 6; };  // 6x6 cells = 18x18 points

constexpr int kStageThreads = 256;

template <int DOF, int KIND>
__global__ __launch_bounds__(kStageThreads) void stage_kernel(MeshDev m, StageArgs a,
                                                              int tiles_x) {
  constexpr int H = DOF - 1;
  constexpr int C = Tile<DOF>::C;
  constexpr int NR = C * H + 1;   // core extent incl. the far cell end
  constexpr int TL = NR + 2 * H;  // staged extent incl. the halo cell
  __shared__ double sE[TL * TL], sU[TL * TL], sV[TL * TL], sPG[TL * TL], sCU[TL * TL],
      sCV[TL * TL];
  __shared__ double lamx[(C + 1) * NR], lamy[(C + 1) * NR];
  __shared__ double rxF[(C + 1) * NR], rxD[(C + 1) * NR], rxV[(C + 1) * NR];
  __shared__ double ryF[(C + 1) * NR], ryD[(C + 1) * NR], ryU[(C + 1) * NR];

  const int npv = m.npv, nh = m.nh, N = m.n;
  const double dx = m.dx;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int p = L % 6 + 1;
  const int tile = L / 6;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  const int cx0 = tx * C, cy0 = ty * C;
  const int ncx = min(C, N - cx0), ncy = min(C, N - cy0);
  const int i0 = cx0 * H + 1, j0 = cy0 * H + 1;
  const int nx = ncx * H + 1, ny = ncy * H + 1;           // core incl. far cell end
  const int TXo = ncx * H + (cx0 + ncx == N ? 1 : 0);     // owned points
  const int TYo = ncy * H + (cy0 + ncy == N ? 1 : 0);
  const int lw = nx + 2 * H, lh = ny + 2 * H;
  const long long pbase = (long long)(p - 1) * nh * nh;
  const long long mbase = (long long)(p - 1) * m.metric_pstride;
  const long long sbase = (long long)(p - 1) * m.phis_pstride;

  // Phase 0: stage the cross-shaped tile + halo; derived fields per point.
  for (int t = threadIdx.x; t < lw * lh; t += blockDim.x) {
    const int ly = t / lw, lx = t - ly * lw;
    const bool inx = lx >= H && lx < H + nx, iny = ly >= H && ly < H + ny;
    if (!inx && !iny) continue;  // halo corner: never read
    const long long off = (long long)(j0 - 1 + ly) * nh + (i0 - 1 + lx);
    const long long g = pbase + off;
    const double phi = a.in_phi[g], u = a.in_u[g], v = a.in_v[g];
    const double sg = m.sqrtG[mbase + off];
    const double g11 = m.ig11[mbase + off], g12 = m.ig12[mbase + off],
                 g22 = m.ig22[mbase + off];
    const double ps = m.phi_s[sbase + off];
    double cu, cv;
    if (a.cu0) {
      cu = a.cu0[g];
      cv = a.cv0[g];
    } else {  // cov2contrav (projection_mod.F90:138-148)
      cu = g11 * u + g12 * v;
      cv = g12 * u + g22 * v;
    }
    const double K = 0.5 * (cu * u + cv * v);  // spatial_operators_mod.F90:60-61
    const int li = ly * TL + lx;
    sE[li] = phi + ps + K;
    sU[li] = u;
    sV[li] = v;
    sPG[li] = phi * sg;
    sCU[li] = cu;
    sCV[li] = cv;
    if (inx && iny) {
      const int ox = lx - H, oy = ly - H;
      if (ox % H == 0) lamx[(ox / H) * NR + oy] = wave_speed(cu, phi, g11);
      if (oy % H == 0) lamy[(oy / H) * NR + ox] = wave_speed(cv, phi, g22);
    }
  }
  __syncthreads();

  // Phase 1: Riemann derivatives at every cell end of the core, x then y.
  const int tasks_x = (ncx + 1) * ny, tasks_y = (ncy + 1) * nx;
  for (int t = threadIdx.x; t < tasks_x + tasks_y; t += blockDim.x) {
    const bool isx = t < tasks_x;
    const int tt = isx ? t : t - tasks_x;
    const int span = isx ? ny : nx;
    const int mb = tt / span, o = tt - mb * span;
    const int st = isx ? 1 : TL;  // stride along the sweep
    const int b = isx ? (o + H) * TL + (H + mb * H) : (H + mb * H) * TL + (o + H);
    const double *sC = isx ? sCU : sCV;
    const double *sQ = isx ? sU : sV;   // q of the gradient flux
    const double *sW = isx ? sV : sU;   // upwinded cross wind
    double eL[DOF], eR[DOF], qL[DOF], qR[DOF], dL[DOF], dR[DOF], pL[DOF], pR[DOF], wL[DOF],
        wR[DOF];
#pragma unroll
    for (int k = 0; k < DOF; ++k) {
      const int l = b - H * st + k * st, r = b + k * st;
      eL[k] = sE[l];  eR[k] = sE[r];
      qL[k] = sQ[l];  qR[k] = sQ[r];
      pL[k] = sPG[l]; pR[k] = sPG[r];
      dL[k] = sPG[l] * sC[l];
      dR[k] = sPG[r] * sC[r];
      wL[k] = sW[l];  wR[k] = sW[r];
    }
    const double lam = isx ? lamx[mb * NR + o] : lamy[mb * NR + o];
    const double up = upwind_sign(sC[b]);
    const double fF = riemann(fit_right<DOF>(eL, dx), fit_left<DOF>(eR, dx),
                              fit_right<DOF>(qL, dx), fit_left<DOF>(qR, dx), lam);
    const double fD = riemann(fit_right<DOF>(dL, dx), fit_left<DOF>(dR, dx),
                              fit_right<DOF>(pL, dx), fit_left<DOF>(pR, dx), lam);
    const double wlr = fit_right<DOF>(wL, dx), wrl = fit_left<DOF>(wR, dx);
    const double fW = riemann(wlr, wrl, wlr, wrl, up);
    if (isx) {
      rxF[mb * NR + o] = fF; rxD[mb * NR + o] = fD; rxV[mb * NR + o] = fW;
    } else {
      ryF[mb * NR + o] = fF; ryD[mb * NR + o] = fD; ryU[mb * NR + o] = fW;
    }
  }
  __syncthreads();

  // Phase 2: derivatives at owned points, tendencies, RK update.
  for (int t = threadIdx.x; t < TXo * TYo; t += blockDim.x) {
    const int oy = t / TXo, ox = t - oy * TXo;
    const int lx = ox + H, ly = oy + H, li = ly * TL + lx;
    double fx, dvx, vx, fy, dvy, uy;
    {  // x sweep
      const int mb = ox / H, pos = ox - mb * H;
      if (pos == 0) {
        fx = rxF[mb * NR + oy]; dvx = rxD[mb * NR + oy]; vx = rxV[mb * NR + oy];
      } else {
        const int b0 = ly * TL + H + mb * H;
        double fe[DOF], fd[DOF], fv[DOF];
#pragma unroll
        for (int k = 0; k < DOF; ++k) {
          fe[k] = sE[b0 + k]; fd[k] = sPG[b0 + k] * sCU[b0 + k]; fv[k] = sV[b0 + k];
        }
        const int r0 = mb * NR + oy, r1 = (mb + 1) * NR + oy;
        fx = cell_interior<DOF>(fe, rxF[r0], rxF[r1], pos, dx);
        dvx = cell_interior<DOF>(fd, rxD[r0], rxD[r1], pos, dx);
        vx = cell_interior<DOF>(fv, rxV[r0], rxV[r1], pos, dx);
      }
    }
    {  // y sweep
      const int mb = oy / H, pos = oy - mb * H;
      if (pos == 0) {
        fy = ryF[mb * NR + ox]; dvy = ryD[mb * NR + ox]; uy = ryU[mb * NR + ox];
      } else {
        const int b0 = (H + mb * H) * TL + lx;
        double fe[DOF], fd[DOF], fu[DOF];
#pragma unroll
        for (int k = 0; k < DOF; ++k) {
          const int l = b0 + k * TL;
          fe[k] = sE[l]; fd[k] = sPG[l] * sCV[l]; fu[k] = sU[l];
        }
        const int r0 = mb * NR + ox, r1 = (mb + 1) * NR + ox;
        fy = cell_interior<DOF>(fe, ryF[r0], ryF[r1], pos, dx);
        dvy = cell_interior<DOF>(fd, ryD[r0], ryD[r1], pos, dx);
        uy = cell_interior<DOF>(fu, ryU[r0], ryU[r1], pos, dx);
      }
    }
    const int i = i0 + ox, j = j0 + oy;
    const long long off = (long long)(j + H - 1) * nh + (i + H - 1);
    const long long g = pbase + off;
    const double sg = m.sqrtG[mbase + off];
    const double fc = m.f[pbase + off];
    const double cu = sCU[li], cv = sCV[li];
    // spatial_operators_mod.F90:127-131
    const double vort = (vx - uy) / sg;
    const double kpg = -dvx - dvy;
    const double ku = -fx + sg * (vort + fc) * cv;
    const double kv = -fy - sg * (vort + fc) * cu;

    if constexpr (KIND == kTend) {
      const long long ti = ((long long)(p - 1) * npv + (j - 1)) * npv + (i - 1);
      a.o_phi[ti] = kpg;
      a.o_u[ti] = ku;
      a.o_v[ti] = kv;
      continue;
    } else {
      double pgn, un, vn;  // base state Y_n
      if constexpr (KIND == kRK4S1 || KIND == kRK3S1) {
        pgn = sPG[li]; un = sU[li]; vn = sV[li];
      } else {
        pgn = a.b_phi[g] * sg; un = a.b_u[g]; vn = a.b_v[g];
      }
      double pg, u, v;
      if constexpr (KIND == kRK4S1) {
        pg = pgn + a.c1 * kpg; u = un + a.c1 * ku; v = vn + a.c1 * kv;
        a.acc_phiG[g] = kpg; a.acc_u[g] = ku; a.acc_v[g] = kv;
      } else if constexpr (KIND == kRK4S2 || KIND == kRK4S3) {
        pg = pgn + a.c1 * kpg; u = un + a.c1 * ku; v = vn + a.c1 * kv;
        a.acc_phiG[g] = a.acc_phiG[g] + 2.0 * kpg;
        a.acc_u[g] = a.acc_u[g] + 2.0 * ku;
        a.acc_v[g] = a.acc_v[g] + 2.0 * kv;
      } else if constexpr (KIND == kRK4S4) {
        const double tpg = (a.acc_phiG[g] + kpg) / 6.0;
        const double tu = (a.acc_u[g] + ku) / 6.0;
        const double tv = (a.acc_v[g] + kv) / 6.0;
        pg = pgn + a.c1 * tpg; u = un + a.c1 * tu; v = vn + a.c1 * tv;
      } else if constexpr (KIND == kRK3S1) {
        pg = pgn + a.c1 * kpg; u = un + a.c1 * ku; v = vn + a.c1 * kv;
      } else if constexpr (KIND == kRK3S2) {
        pg = 0.75 * pgn + 0.25 * sPG[li] + a.c1 * kpg;
        u = 0.75 * un + 0.25 * sU[li] + a.c1 * ku;
        v = 0.75 * vn + 0.25 * sV[li] + a.c1 * kv;
      } else {  // kRK3S3
        pg = pgn / 3.0 + 2.0 / 3.0 * sPG[li] + a.c1 * kpg;
        u = un / 3.0 + 2.0 / 3.0 * sU[li] + a.c1 * ku;
        v = vn / 3.0 + 2.0 / 3.0 * sV[li] + a.c1 * kv;
      }
      const double phi = pg / sg;  // time_scheme_mod.F90:81
      a.o_phi[g] = phi;
      a.o_u[g] = u;
      a.o_v[g] = v;
      // raw copy of edge-line points for the boundary pass
      if (i == 1 || i == npv || j == 1 || j == npv) {
        const double vals[3] = {phi, u, v};
#pragma unroll
        for (int fl = 0; fl < 3; ++fl) {
          if (i == 1) a.edge_raw[edge_slot(npv, p, 0, fl, j)] = vals[fl];
          if (i == npv) a.edge_raw[edge_slot(npv, p, 1, fl, j)] = vals[fl];
          if (j == 1) a.edge_raw[edge_slot(npv, p, 2, fl, i)] = vals[fl];
          if (j == npv) a.edge_raw[edge_slot(npv, p, 3, fl, i)] = vals[fl];
        }
      }
    }
  }
}

// Raw (pre-averaging) phi, u, v of interior point (q, i, j) after a stage:
// edge-line points come from the raw copy, the rest from the level itself.
__device__ __forceinline__ void raw_point(const MeshDev &m, const BoundaryArgs &a, int q, int i,
                                          int j, double &phi, double &u, double &v) {
  const int npv = m.npv;
  int side = -1, s = 0;
  if (i == 1) { side = 0; s = j; }
  else if (i == npv) { side = 1; s = j; }
  else if (j == 1) { side = 2; s = i; }
  else if (j == npv) { side = 3; s = i; }
  if (side >= 0) {
    phi = a.edge_raw[edge_slot(npv, q, side, 0, s)];
    u = a.edge_raw[edge_slot(npv, q, side, 1, s)];
    v = a.edge_raw[edge_slot(npv, q, side, 2, s)];
  } else {
    const long long g = pidx(m.nh, m.dof - 1, q, i, j);
    phi = a.phi[g];
    u = a.u[g];
    v = a.v[g];
  }
}

__device__ __forceinline__ void edge_ij(int npv, int side, int s, int &i, int &j) {
  if (side == 0) { i = 1; j = s; }
  else if (side == 1) { i = npv; j = s; }
  else if (side == 2) { i = s; j = 1; }
  else { i = s; j = npv; }
}

// correct_bdy_ghost (time_scheme_mod.F90:125-142) on one level: one thread per
// (panel, side, s), plus 8 threads for the cube vertices.
template <int DOF>
__global__ __launch_bounds__(256) void boundary_kernel(MeshDev m, BoundaryArgs a) {
  constexpr int H = DOF - 1;
  const int npv = m.npv, nh = m.nh;
  const double dx = m.dx;
  const int nside = 6 * 4 * npv;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nside + 8;
       t += gridDim.x * blockDim.x) {
  if (t < nside) {
    const int s = t % npv + 1, ps = t / npv, side = ps % 4, p = ps / 4 + 1;
    const HaloSrc hs = kHalo[p - 1][side];
    // Ghost points at depths 1..H (ghost_mod.f90:154-227), sphere winds
    // through P2SP at the neighbour's points and back through SP2P at the
    // ghost point (spatial_operators_mod.F90:408-492).
    for (int k = 1; k <= H; ++k) {
      const int c = m.giref[(k - 1) * npv + (s - 1)];
      const double cf = m.gcoef[(k - 1) * npv + (s - 1)];
      const int nrm = hs.far ? npv - k : k + 1;
      double qp[DOF], qz[DOF], qm[DOF];
#pragma unroll
      for (int l = 0; l < DOF; ++l) {
        const int tpos = (c - 1) * H + 1 + l;
        const int al = hs.rev ? npv + 1 - tpos : tpos;
        const int qi = hs.col ? nrm : al, qj = hs.col ? al : nrm;
        double phi, u, v;
        raw_point(m, a, hs.q, qi, qj, phi, u, v);
        qp[l] = phi;
        apply2(m.p2sp + 4 * pidx(nh, H, hs.q, qi, qj), u, v, qz[l], qm[l]);
      }
      int hi, hj;
      if (side == 0) { hi = 1 - k; hj = s; }
      else if (side == 1) { hi = npv + k; hj = s; }
      else if (side == 2) { hi = s; hj = 1 - k; }
      else { hi = s; hj = npv + k; }
      const long long g = pidx(nh, H, p, hi, hj);
      const double zh = ghost_interp<DOF>(qz, cf, dx), mh = ghost_interp<DOF>(qm, cf, dx);
      a.phi[g] = ghost_interp<DOF>(qp, cf, dx);
      double uh, vh;
      apply2(m.sp2p + 4 * g, zh, mh, uh, vh);
      a.u[g] = uh;
      a.v[g] = vh;
    }
    // Edge averaging (spatial_operators_mod.F90:298-345), vertices excluded.
    const int e = kFirstEdge[p - 1][side];
    if (e >= 0 && s > 1 && s < npv) {
      const EdgePair ep = kEdges[e];
      const int sb = ep.rev ? npv + 1 - s : s;
      const double pa = a.edge_raw[edge_slot(npv, ep.pa, ep.sa, 0, s)];
      const double ua = a.edge_raw[edge_slot(npv, ep.pa, ep.sa, 1, s)];
      const double va = a.edge_raw[edge_slot(npv, ep.pa, ep.sa, 2, s)];
      const double pb = a.edge_raw[edge_slot(npv, ep.pb, ep.sb, 0, sb)];
      const double ub = a.edge_raw[edge_slot(npv, ep.pb, ep.sb, 1, sb)];
      const double vb = a.edge_raw[edge_slot(npv, ep.pb, ep.sb, 2, sb)];
      int ia, ja, ib, jb;
      edge_ij(npv, ep.sa, s, ia, ja);
      edge_ij(npv, ep.sb, sb, ib, jb);
      const long long ga = pidx(nh, H, ep.pa, ia, ja), gb = pidx(nh, H, ep.pb, ib, jb);
      double za, ma, zb, mb;
      apply2(m.p2sp + 4 * ga, ua, va, za, ma);
      apply2(m.p2sp + 4 * gb, ub, vb, zb, mb);
      const double ph = 0.5 * (pa + pb), zu = 0.5 * (za + zb), mu = 0.5 * (ma + mb);
      double u1, v1, u2, v2;
      apply2(m.sp2p + 4 * ga, zu, mu, u1, v1);
      apply2(m.sp2p + 4 * gb, zu, mu, u2, v2);
      a.phi[ga] = ph; a.u[ga] = u1; a.v[ga] = v1;
      a.phi[gb] = ph; a.u[gb] = u2; a.v[gb] = v2;
    }
  } else {
    // Cube vertices (spatial_operators_mod.F90:347-386), reference summation order.
    const int vx = t - nside;
    double ph[3], z[3], mr[3];
    long long gg[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const VtxPt vp = kVerts[vx][l];
      const int i = vp.ci ? npv : 1, j = vp.cj ? npv : 1;
      const int side = vp.ci ? 1 : 0;  // any edge through the corner holds its raw copy
      ph[l] = a.edge_raw[edge_slot(npv, vp.p, side, 0, j)];
      const double u = a.edge_raw[edge_slot(npv, vp.p, side, 1, j)];
      const double v = a.edge_raw[edge_slot(npv, vp.p, side, 2, j)];
      gg[l] = pidx(nh, H, vp.p, i, j);
      apply2(m.p2sp + 4 * gg[l], u, v, z[l], mr[l]);
    }
    const double pv = (ph[0] + ph[1] + ph[2]) / 3.0;
    const double zv = (z[0] + z[1] + z[2]) / 3.0;
    const double mv = (mr[0] + mr[1] + mr[2]) / 3.0;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      double u, v;
      apply2(m.sp2p + 4 * gg[l], zv, mv, u, v);
      a.phi[gg[l]] = pv;
      a.u[gg[l]] = u;
      a.v[gg[l]] = v;
    }
  }
  }
}

// Reference-layout copy-out of one level: phi, phiG = phi*sqrtG, u, v and
// the contravariant winds (or the host's initial ones before the first step).
__global__ void pack_kernel(MeshDev m, const double *phi, const double *u, const double *v,
                            const double *cu0, const double *cv0, double *o_phi,
                            double *o_phiG, double *o_u, double *o_v, double *o_cu,
                            double *o_cv) {
  const long long np = 6LL * m.nh * m.nh;
  const long long per = (long long)m.nh * m.nh;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < np;
       g += (long long)gridDim.x * blockDim.x) {
    const long long p = g / per, off = g - p * per;
    const long long mi = p * m.metric_pstride + off;
    const double ph = phi[g], uu = u[g], vv = v[g];
    o_phi[g] = ph;
    o_phiG[g] = ph * m.sqrtG[mi];
    o_u[g] = uu;
    o_v[g] = vv;
    if (cu0) {
      o_cu[g] = cu0[g];
      o_cv[g] = cv0[g];
    } else {
      o_cu[g] = m.ig11[mi] * uu + m.ig12[mi] * vv;
      o_cv[g] = m.ig12[mi] * uu + m.ig22[mi] * vv;
    }
  }
}

// Sphere winds on every point (spatial_operators_mod.F90:391-405).
__global__ void sphere_winds_kernel(MeshDev m, const double *u, const double *v, double *zon,
                                    double *mer) {
  const long long np = 6LL * m.nh * m.nh;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < np;
       g += (long long)gridDim.x * blockDim.x) {
    double z, w;
    apply2(m.p2sp + 4 * g, u[g], v[g], z, w);
    zon[g] = z;
    mer[g] = w;
  }
}

// ---------------------------------------------------------------------------
// FP64 integrals (diag_mod.F90:10-67): per-cell sums in the reference order,
// then a deterministic double-double tree over cells and blocks.
// ---------------------------------------------------------------------------
struct DD { double hi, lo; };

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  // two-sum of the high parts, then fold the low parts in (Knuth)
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double err = (a.hi - (s - bb)) + (b.hi - bb);
  const double lo = err + a.lo + b.lo;
  const double hi = s + lo;
  return DD{hi, lo - (hi - s)};
}

constexpr int kDiagThreads = 256;

template <int DOF>
__global__ __launch_bounds__(kDiagThreads) void diag_partial_kernel(MeshDev m, const double *phi,
                                                                    const double *u,
                                                                    const double *v,
                                                                    const double *cu0,
                                                                    const double *cv0,
                                                                    double *partials) {
  constexpr int H = DOF - 1;
  const int N = m.n, nh = m.nh;
  const long long ncell = 6LL * N * N;
  double w1[DOF];
  if constexpr (DOF == 3) { w1[0] = 1.0; w1[1] = 4.0; w1[2] = 1.0; }
  else { w1[0] = 1.0; w1[1] = 3.0; w1[2] = 3.0; w1[3] = 1.0; }
  const double norm = DOF == 3 ? 36.0 : 80.0;  // mesh_mod.F90:183-203
  DD ms{0.0, 0.0}, es{0.0, 0.0};
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(c / ((long long)N * N)) + 1;
    const int rem = (int)(c - (long long)(p - 1) * N * N);
    const int cj = rem / N + 1, ci = rem - (cj - 1) * N + 1;
    double mc = 0.0, ec = 0.0;
    for (int b = 0; b < DOF; ++b)
      for (int aa = 0; aa < DOF; ++aa) {
        const int i = (ci - 1) * H + 1 + aa, j = (cj - 1) * H + 1 + b;
        const long long g = pidx(nh, H, p, i, j);
        const long long off = g - (long long)(p - 1) * nh * nh;
        const long long mi = (long long)(p - 1) * m.metric_pstride + off;
        const double w = (w1[aa] * w1[b]) / norm;
        const double sg = m.sqrtG[mi];
        const double ph = phi[g], uu = u[g], vv = v[g];
        const double cu = cu0 ? cu0[g] : m.ig11[mi] * uu + m.ig12[mi] * vv;
        const double cv = cv0 ? cv0[g] : m.ig12[mi] * uu + m.ig22[mi] * vv;
        const double ps = m.phi_s[(long long)(p - 1) * m.phis_pstride + off];
        mc = mc + w * (ph * sg);
        const double KE = 0.5 * ph * (uu * cu + vv * cv);
        const double PE = 0.5 * ((ph + ps) * (ph + ps));
        ec = ec + w * sg * (KE + PE);
      }
    ms = dd_add(ms, DD{mc, 0.0});
    es = dd_add(es, DD{ec, 0.0});
  }
  __shared__ DD sm[kDiagThreads], se[kDiagThreads];
  sm[threadIdx.x] = ms;
  se[threadIdx.x] = es;
  __syncthreads();
  for (int w = (int)blockDim.x / 2; w > 0; w >>= 1) {  // blockDim.x: a power of two <= kDiagThreads
    if ((int)threadIdx.x < w) {
      sm[threadIdx.x] = dd_add(sm[threadIdx.x], sm[threadIdx.x + w]);
      se[threadIdx.x] = dd_add(se[threadIdx.x], se[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[4 * blockIdx.x + 0] = sm[0].hi;
    partials[4 * blockIdx.x + 1] = sm[0].lo;
    partials[4 * blockIdx.x + 2] = se[0].hi;
    partials[4 * blockIdx.x + 3] = se[0].lo;
  }
}

__global__ void diag_final_kernel(const double *partials, int nblocks, double *out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    DD ms{0.0, 0.0}, es{0.0, 0.0};
    for (int b = 0; b < nblocks; ++b) {
      ms = dd_add(ms, DD{partials[4 * b], partials[4 * b + 1]});
      es = dd_add(es, DD{partials[4 * b + 2], partials[4 * b + 3]});
    }
    out[0] = ms.hi + ms.lo;
    out[1] = es.hi + es.lo;
  }
}

}  // namespace mcv
