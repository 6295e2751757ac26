// Two SGM paths in one sweep of an [H,W,D] float32 volume: an axis path and
// a diagonal path that advance together, so the cost and the accumulator
// are read once for both (post.cu:178-259 for each path's recurrence).
//
// The eight paths pair as {E, SE} and {W, NW} (sweeps along rows: a
// scanline is an image row) and {S, SW} and {N, NE} (sweeps along columns:
// a scanline is an image column). In each pair the diagonal predecessor of
// a pixel lies on the neighbouring scanline, one step back; in every pair
// that neighbour is on the same side, so scanlines have a fixed dependency
// order (rank): rank r needs rank r - 1's diagonal path values of the
// previous step.
//
// One wavefront per scanline, kPairWaves consecutive ranks per workgroup.
// Inside a workgroup the diagonal values (with the predecessor pixel's
// image values) pass between wavefronts through LDS, double-buffered with
// one barrier per step. Between workgroups the last wavefront of rank block
// b writes its values of every step to an edge buffer, and the first
// wavefront of block b + 1 reads them: each word is stored and loaded with
// agent-scope atomic accesses (coherent across XCDs), and a word that still
// holds kEdgeEmpty (a NaN no path value takes) has not been written yet, so
// no fence or flag is needed; the reader writes kEdgeEmpty back after use,
// which leaves the buffer ready for the next launch. Workgroups take their
// rank block from a ticket counter in launch order, so block b waits only
// for block b - 1, which took its ticket earlier and is therefore resident:
// the sweep cannot deadlock, whatever the occupancy. A wait that exceeds
// `spin_limit` wall-clock ticks (it does not happen in correct operation) is
// counted as a timeout, marks the edge buffer for repair and goes on with
// zeros, and later waits of that wavefront do not wait, so every wavefront
// reaches the end; the last workgroup to finish resets the ticket.
//
// The accumulator gets both paths' values added in double and rounded once.
#pragma once

#include "sgm_scan.hpp"

namespace cbmv {
namespace {

constexpr int kPairWaves = 8;
constexpr uint32_t kEdgeEmpty = 0x7fc5e1a7u;

struct PairSync {
  float* edge;     // [blocks - 1][steps][64 * DPL]
  float* edge_mn;  // [blocks - 1][steps]
  unsigned* ctr;   // [0] ticket, [1] finished blocks, [2] timeouts, [3] repair needed
};

__device__ __forceinline__ uint32_t ld_agent(const float* p) {
  return __hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(float* p, float v) {
  __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// ROWS: scanlines are image rows (E/SE with sa = +1, sd = +1; W/NW with
// sa = sd = -1); otherwise image columns (S/SW: sa = +1, sd = -1; N/NE:
// sa = -1, sd = +1). sa is the step direction along the scanline, sd the
// neighbour side: the diagonal predecessor of scanline l is scanline l - sd.
template <int DPL, int PF, int DIR, bool FIRST, bool LAST, bool FULL, bool ROWS>
__global__ void __launch_bounds__(64 * kPairWaves) k_sgm_pair(
    const float* __restrict__ x0, const float* __restrict__ x1p, const float* __restrict__ in,
    float* __restrict__ out, int H, int W, int D, Penalties pen, float tau_so, int sa, int sd,
    PairSync sync, double scale, long long spin_limit) {
  constexpr int K = kPairWaves;
  constexpr int NDP = 64 * DPL;
  // per wavefront and step: its diagonal path values, their minimum, and
  // its pixel's right-image segment and left-image value (the next
  // scanline's diagonal predecessor's)
  __shared__ float sh_d[2][K][NDP];
  __shared__ float sh_s[2][K][NDP];
  __shared__ float sh_mn[2][K];
  __shared__ float sh_x0[2][K];
  __shared__ int sh_block;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int L = ROWS ? H : W;  // scanlines
  const int S = ROWS ? W : H;  // steps per scanline
  const int nblocks = (L + K - 1) / K;
  if (threadIdx.x == 0) sh_block = (int)atomicAdd(sync.ctr + opaque_zero(), 1u);
  __syncthreads();
  // (a ticket past the last block only follows an aborted launch; the
  // modulus keeps every access inside the buffers)
  const int blk = __builtin_amdgcn_readfirstlane(sh_block % nblocks);
  const int rank = blk * K + wv;
  const bool act = rank < L;
  const int line = act ? (sd > 0 ? rank : L - 1 - rank) : 0;
  const bool has_nb = rank >= 1;  // a neighbouring scanline precedes this one
  const bool from_edge = has_nb && wv == 0;
  // the next block's first scanline reads steps 0 .. S-2 (its steps 1 .. S-1)
  const bool to_edge = act && wv == K - 1 && blk < nblocks - 1;
  const int dxa = ROWS ? sa : 0, dya = ROWS ? 0 : sa;
  const int dxd = ROWS ? sa : sd, dyd = ROWS ? sd : sa;

  const int dbase = lane * DPL;
  const int64_t pitch = (int64_t)W + 2 * kSgmPad;
  const int64_t sy = (int64_t)W * D, sx = D;
  int joff[DPL];
  bool jok[DPL];
#pragma unroll
  for (int j = 0; j < DPL; j++) {
    jok[j] = FULL || dbase + j < D;
    joff[j] = FULL ? dbase + j : min(dbase + j, D - 1);
  }
  // the right-image pair (x + d DIR, x + d DIR - dx) is inside the image
  // (post.cu:229-234) iff 0 <= x + d DIR - max(0, dx) < W - |dx|: with
  // xo = dbase DIR - max(0, dx), iff (unsigned)(x + xo + j DIR) < W - |dx|
  const int xoa = dbase * DIR - max(0, dxa), xod = dbase * DIR - max(0, dxd);
  const uint32_t xwa = (uint32_t)max(W - abs(dxa), 0), xwd = (uint32_t)max(W - abs(dxd), 0);
  const int seg0 = DIR > 0 ? dbase : -(dbase + DPL - 1);
  constexpr int kShift = DIR > 0 ? 0 : 64 * DPL;
  const uint32_t lane_v = (uint32_t)(dbase * 4);
  const uint32_t lane_p = (uint32_t)((seg0 + kShift) * 4);
  const int lane_0 = opaque_zero();

  // pixel of step s on this scanline
  auto px = [&](int s) { return ROWS ? (sa > 0 ? s : W - 1 - s) : line; };
  auto py = [&](int s) { return ROWS ? line : (sa > 0 ? s : H - 1 - s); };

  float rin[PF][DPL], rout[PF][DPL], rseg[PF][DPL], rx0[PF];
  // first wavefront: the previous block's values (edge buffer) and the
  // predecessor pixel's image values, PF - 1 steps ahead
  float er[PF][DPL], emn[PF], esg[PF][DPL], ex0[PF];

  auto fetch = [&](int slot, int s) {
    s = min(s, S - 1);  // past the end: repeat the last step (never used)
    const int x = px(s), y = py(s);
    if (act) {
      const float* pin = in + y * sy + x * sx;
      const float* pacc = out + y * sy + x * sx;
      if (FULL) {
        vload((const float*)((const char*)pin + lane_v), rin[slot]);
        if (!FIRST) vload((const float*)((const char*)pacc + lane_v), rout[slot]);
      } else {
#pragma unroll
        for (int j = 0; j < DPL; j++) {
          rin[slot][j] = pin[joff[j]];
          if (!FIRST) rout[slot][j] = pacc[joff[j]];
        }
      }
      const char* p1 = (const char*)x1p + (y * pitch + kSgmPad + x - kShift) * 4;
      vload((const float*)(p1 + lane_p), rseg[slot]);
      rx0[slot] = x0[((int64_t)y * W + x) + lane_0];
    }
  };
  // edge entry of step s (the previous block's last scanline at step s - 1)
  auto edge_at = [&](int s) { return (int64_t)(blk - 1) * S + (s - 1); };
  auto efetch = [&](int slot, int s) {
    if (from_edge && s >= 1 && s < S) {
      const int64_t e = edge_at(s);
#pragma unroll
      for (int j = 0; j < DPL; j++)
        er[slot][j] = __uint_as_float(ld_agent(sync.edge + e * NDP + dbase + j));
      emn[slot] = __uint_as_float(ld_agent(sync.edge_mn + e + lane_0));
      // the diagonal predecessor pixel, on the previous block's last scanline
      const int xd = px(s) - dxd, yd = py(s) - dyd;
      const char* p1 = (const char*)x1p + (yd * pitch + kSgmPad + xd - kShift) * 4;
      vload((const float*)(p1 + lane_p), esg[slot]);
      ex0[slot] = x0[((int64_t)yd * W + xd) + lane_0];
    }
  };

#pragma unroll
  for (int k = 0; k < PF - 1; k++) {
    fetch(k, k);
    efetch(k, k);
  }
#pragma unroll
  for (int j = 0; j < DPL; j++) rseg[PF - 1][j] = 0.f;
  rx0[PF - 1] = 0.f;

  // penalty tables, one VGPR per path: value i in lane i, [0..2] for the
  // d-1 neighbour per class, [3..5] for d+1, [6..8] P2 (post.cu:239-255);
  // P1/alpha1 on d-1 for downward paths, on d+1 for upward ones
  auto table = [&](int dy) {
    const bool down = dy > 0, up = dy < 0;
    const float v[9] = {down ? pen.p1a[0] : pen.p1[0], down ? pen.p1a[1] : pen.p1[1],
                        down ? pen.p1a[2] : pen.p1[2], up ? pen.p1a[0] : pen.p1[0],
                        up ? pen.p1a[1] : pen.p1[1],   up ? pen.p1a[2] : pen.p1[2],
                        pen.p2[0],                     pen.p2[1],
                        pen.p2[2]};
    float t = 0.f;
#pragma unroll
    for (int i = 0; i < 9; i++) t = lane == i ? v[i] : t;
    return t;
  };
  const float ptab_a = table(dya), ptab_d = table(dyd);
  auto rl = [&](float t, int i) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), i));
  };
  const float Mla = in_vgpr(rl(ptab_a, 2)), Mra = in_vgpr(rl(ptab_a, 5)),
              M2a = in_vgpr(rl(ptab_a, 8));
  const float Mld = in_vgpr(rl(ptab_d, 2)), Mrd = in_vgpr(rl(ptab_d, 5)),
              M2d = in_vgpr(rl(ptab_d, 8));
  const float ten = in_vgpr(10.f), inf = in_vgpr(INFINITY);
  const int taui = __float_as_int(tau_so + 0.f);
  float prev[DPL];
#pragma unroll
  for (int j = 0; j < DPL; j++) prev[j] = 0.f;
  float mn = 0.f;
  // after one wait has given up, later ones do not wait: a sweep stalls for
  // at most one spin_limit per wavefront
  bool gave_up = false;

  // one path's new values into acc and cur: pv/mnp the predecessor's
  // values and minimum, (seg, pseg) the right-image pair, d1 the left-image
  // difference
  auto path_vals = [&](const float (&pv)[DPL], float mnp, const float (&cost)[DPL],
                       const float (&seg)[DPL], const float (&pseg)[DPL], float d1, int x,
                       int xo, uint32_t xw, float ptab, float Ml, float Mr, float M2,
                       double (&acc)[DPL], float (&cur)[DPL]) {
    const float left_nb = __int_as_float(__builtin_amdgcn_update_dpp(
        __float_as_int(inf), __float_as_int(pv[DPL - 1]), 0x138, 0xf, 0xf, false));
    const float right_nb = __int_as_float(__builtin_amdgcn_update_dpp(
        __float_as_int(inf), __float_as_int(pv[0]), 0x130, 0xf, 0xf, false));
    const int d1i = __builtin_amdgcn_readfirstlane(__float_as_int(d1));
    const bool lt1 = d1i < taui, gt1 = d1i > taui;
    const int cls = lt1 ? 0 : (gt1 ? 1 : 2);
    // D2's class test as one compare: sg D2 < tb with (sg, tb) = (1, tau)
    // when D1 < tau, (-1, -tau) when D1 > tau, (0, 0) (never true) otherwise
    const float sg = lt1 ? 1.f : (gt1 ? -1.f : 0.f);
    const float tb = lt1 ? tau_so : (gt1 ? -tau_so : 0.f);
    const float vAl = rl(ptab, cls), vAr = rl(ptab, 3 + cls), vA2 = rl(ptab, 6 + cls);
    const double mnd = (double)mnp;
#pragma unroll
    for (int j = 0; j < DPL; j++) {
      const int d = dbase + j;
      const float sj = DIR > 0 ? seg[j] : seg[DPL - 1 - j];
      const float pj = DIR > 0 ? pseg[j] : pseg[DPL - 1 - j];
      const bool ok2 = (uint32_t)(x + xo + j * DIR) < xw;
      const float D2 = ok2 ? fabsf(sj - pj) : ten;
      const bool cd = sg * D2 < tb;
      const float Pl = cd ? vAl : Ml;
      const float Pr = cd ? vAr : Mr;
      const float P2 = cd ? vA2 : M2;
      const float pl = j > 0 ? pv[j - 1] : left_nb;
      float pr = j < DPL - 1 ? pv[j + 1] : right_nb;
      if (j == DPL - 1 && !FULL) pr = d + 1 < D ? pr : inf;
      const float c = fminf(fminf(pv[j], mnp + P2), fminf(pl + Pl, pr + Pr));
      // a first pixel (no predecessor): pv = 0 and mnp = 0, so with
      // non-negative penalties c = 0 and the value is the cost itself
      const double val = ((double)cost[j] + (double)c) - mnd;
      acc[j] += val;
      cur[j] = jok[j] ? (float)val : INFINITY;
    }
  };

  auto step = [&](int k, int s) {
    __builtin_amdgcn_sched_barrier(0);
    const int kp = (k + PF - 1) % PF;
    const int buf = s & 1;
    const int x = px(s);
    // the diagonal predecessor's values, minimum and image values
    float pd[DPL], psg[DPL];
    float mnd = 0.f, px0 = 0.f;
    const bool pred = has_nb && s >= 1;
    if (!pred) {
#pragma unroll
      for (int j = 0; j < DPL; j++) {
        pd[j] = 0.f;
        psg[j] = rseg[k][j];
      }
      px0 = rx0[k];
    } else if (from_edge) {
      bool empty = __float_as_uint(emn[k]) == kEdgeEmpty;
#pragma unroll
      for (int j = 0; j < DPL; j++) empty |= __float_as_uint(er[k][j]) == kEdgeEmpty;
      if (__builtin_amdgcn_ballot_w64(empty) != 0) {
        // not yet written: poll (bounded) until every word of the entry is
        const int64_t e = edge_at(s);
        const long long t0 = wall_clock64();
        while (!gave_up) {
          __builtin_amdgcn_s_sleep(2);
#pragma unroll
          for (int j = 0; j < DPL; j++)
            er[k][j] = __uint_as_float(ld_agent(sync.edge + e * NDP + dbase + j));
          emn[k] = __uint_as_float(ld_agent(sync.edge_mn + e + lane_0));
          empty = __float_as_uint(emn[k]) == kEdgeEmpty;
#pragma unroll
          for (int j = 0; j < DPL; j++) empty |= __float_as_uint(er[k][j]) == kEdgeEmpty;
          if (__builtin_amdgcn_ballot_w64(empty) == 0) break;
          if (wall_clock64() - t0 > spin_limit) {
            gave_up = true;
            if (lane == 0) {
              atomicAdd(sync.ctr + 2 + lane_0, 1u);
              atomicOr(sync.ctr + 3 + lane_0, 1u);
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < DPL; j++) {
        pd[j] = __float_as_uint(er[k][j]) != kEdgeEmpty ? er[k][j] : 0.f;
        psg[j] = esg[k][j];
      }
      mnd = __float_as_uint(emn[k]) != kEdgeEmpty ? emn[k] : 0.f;
      px0 = ex0[k];
      // the entry is consumed: empty again for the next launch
      const int64_t e = edge_at(s);
#pragma unroll
      for (int j = 0; j < DPL; j++) sync.edge[e * NDP + dbase + j] = __uint_as_float(kEdgeEmpty);
      if (lane == 0) sync.edge_mn[e] = __uint_as_float(kEdgeEmpty);
    } else {
#pragma unroll
      for (int j = 0; j < DPL; j++) {
        pd[j] = sh_d[buf ^ 1][wv - 1][dbase + j];
        psg[j] = sh_s[buf ^ 1][wv - 1][dbase + j];
      }
      mnd = sh_mn[buf ^ 1][wv - 1];
      px0 = sh_x0[buf ^ 1][wv - 1];
    }
    double acc[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++) acc[j] = FIRST ? 0.0 : (double)rout[k][j];
    float ca[DPL], cdg[DPL];
    path_vals(prev, mn, rin[k], rseg[k], rseg[kp], fabsf(rx0[k] - rx0[kp]), x, xoa, xwa, ptab_a,
              Mla, Mra, M2a, acc, ca);
    path_vals(pd, mnd, rin[k], rseg[k], psg, fabsf(rx0[k] - px0), x, xod, xwd, ptab_d, Mld, Mrd,
              M2d, acc, cdg);
    float la = INFINITY, ld = INFINITY;
    float o[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++) {
      o[j] = (float)(LAST ? acc[j] * scale : acc[j]);
      la = fminf(la, ca[j]);
      ld = fminf(ld, cdg[j]);
      prev[j] = ca[j];
    }
    mn = wave_min_dpp(la);
    const float mdo = wave_min_dpp(ld);
    if (act) {
      float* po = out + py(s) * sy + x * sx;
      if (FULL) {
        vstore((float*)((char*)po + lane_v), o);
      } else {
#pragma unroll
        for (int j = 0; j < DPL; j++)
          if (jok[j]) po[joff[j]] = o[j];
      }
#pragma unroll
      for (int j = 0; j < DPL; j++) {
        sh_d[buf][wv][dbase + j] = cdg[j];
        sh_s[buf][wv][dbase + j] = rseg[k][j];
      }
      if (lane == 0) {
        sh_mn[buf][wv] = mdo;
        sh_x0[buf][wv] = rx0[k];
      }
      if (to_edge && s < S - 1) {
        const int64_t e = (int64_t)blk * S + s;
#pragma unroll
        for (int j = 0; j < DPL; j++) st_agent(sync.edge + e * NDP + dbase + j, cdg[j]);
        if (lane == 0) st_agent(sync.edge_mn + e + lane_0, mdo);
      }
    }
    __syncthreads();
  };

  int s0 = 0;
  for (; s0 + PF <= S; s0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      step(k, s0 + k);
      fetch((k + PF - 1) % PF, s0 + k + PF - 1);
      efetch((k + PF - 1) % PF, s0 + k + PF - 1);
    }
  }
#pragma unroll
  for (int k = 0; k < PF; k++) {
    if (s0 + k < S) {
      step(k, s0 + k);
      fetch((k + PF - 1) % PF, s0 + k + PF - 1);
      efetch((k + PF - 1) % PF, s0 + k + PF - 1);
    }
  }
  // the last block to finish resets the ticket for the next launch
  if (threadIdx.x == 0) {
    const unsigned done = atomicAdd(sync.ctr + 1 + lane_0, 1u);
    if (done == (unsigned)nblocks - 1) {
      __hip_atomic_store(sync.ctr + 0 + lane_0, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(sync.ctr + 1 + lane_0, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Sets every word of the edge buffer to kEdgeEmpty: at set-up (always =
// true), and before each post-processing chain when a sweep timed out
// (ctr[3] set; does nothing otherwise). k_pair_repaired then clears ctr[3].
__global__ void k_pair_fill(PairSync sync, int64_t n_edge, int64_t n_mn, bool always) {
  if (!always && __hip_atomic_load(sync.ctr + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_edge;
       i += (int64_t)gridDim.x * blockDim.x)
    sync.edge[i] = __uint_as_float(kEdgeEmpty);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_mn;
       i += (int64_t)gridDim.x * blockDim.x)
    sync.edge_mn[i] = __uint_as_float(kEdgeEmpty);
}

__global__ void k_pair_repaired(PairSync sync) {
  if (threadIdx.x == 0)
    __hip_atomic_store(sync.ctr + 3, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace
}  // namespace cbmv
