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
// One wavefront per scanline, pair_waves<ROWS>() consecutive ranks per workgroup.
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
// Each step is float32 arithmetic: a path's value is cost + (c - min)
// (the reference evaluates cost + c - min in double and stores the path
// value as float; here two roundings instead of one), and the accumulator
// gets both paths' values added in float. The minimum over d of the
// non-negative path values is an unsigned minimum over their bits, one
// DPP-combined v_min_u32 per reduction step.
#pragma once

#include "sgm_scan.hpp"

namespace cbmv {
namespace {

// wavefronts (scanlines) per workgroup: 8 for the row sweeps (1000 rows at
// 1500x1000: 125 workgroups), 12 for the column sweeps (125 instead of 188
// workgroups, so fewer hand-overs; their loop fits the 168 VGPRs of three
// wavefronts per SIMD with three steps of prefetch)
template <bool ROWS>
constexpr int pair_waves() {
  return ROWS ? 8 : 12;
}
constexpr uint32_t kEdgeEmpty = 0x7fc5e1a7u;

struct PairSync {
  float* edge;     // [blocks - 1][steps][64 * DPL]
  float* edge_mn;  // [blocks - 1][steps]
  unsigned* ctr;   // [0] ticket, [1] finished blocks, [2] timeouts, [3] repair needed
};

// N floats at byte offset voff of a buffer descriptor (16-byte pieces for
// N a multiple of 4), and the store of N floats
template <int N>
__device__ __forceinline__ void bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, float (&v)[N]) {
  if constexpr (N % 4 == 0) {
#pragma unroll
    for (int q = 0; q < N / 4; q++) {
      const auto t = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(voff + 16 * q), 0, 0);
#pragma unroll
      for (int i = 0; i < 4; i++) v[4 * q + i] = __uint_as_float(t[i]);
    }
  } else if constexpr (N == 2) {
    const auto t = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, 0, 0);
    v[0] = __uint_as_float(t[0]);
    v[1] = __uint_as_float(t[1]);
  } else {
    static_assert(N == 1, "1, 2 or a multiple of 4 floats");
    v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, 0));
  }
}
template <int N>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, const float (&v)[N]) {
  if constexpr (N % 4 == 0) {
#pragma unroll
    for (int q = 0; q < N / 4; q++) {
      typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
      const u32x4_t t = {__float_as_uint(v[4 * q]), __float_as_uint(v[4 * q + 1]),
                         __float_as_uint(v[4 * q + 2]), __float_as_uint(v[4 * q + 3])};
      __builtin_amdgcn_raw_buffer_store_b128(t, r, (int)(voff + 16 * q), 0, 0);
    }
  } else if constexpr (N == 2) {
    typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
    const u32x2_t t = {__float_as_uint(v[0]), __float_as_uint(v[1])};
    __builtin_amdgcn_raw_buffer_store_b64(t, r, (int)voff, 0, 0);
  } else {
    static_assert(N == 1, "1, 2 or a multiple of 4 floats");
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[0]), r, (int)voff, 0, 0);
  }
}
// descriptor of the bytes from p on (no bound that an access could reach)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(const char* p) {
  return buffer_rsrc(p, 0x7fffffffu);
}

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
__global__ void __launch_bounds__(64 * pair_waves<ROWS>()) k_sgm_pair(
    const float* __restrict__ x0, const float* __restrict__ x1p, const float* __restrict__ in,
    float* __restrict__ out, int H, int W, int D, Penalties pen, float tau_so, int sa, int sd,
    PairSync sync, double scale, long long spin_limit) {
  constexpr int K = pair_waves<ROWS>();
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
  const int dya = ROWS ? 0 : sa;
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
  const int seg0 = DIR > 0 ? dbase : -(dbase + DPL - 1);
  constexpr int kShift = DIR > 0 ? 0 : 64 * DPL;
  const uint32_t lane_v = (uint32_t)(dbase * 4);
  const uint32_t lane_p = (uint32_t)((seg0 + kShift) * 4);
  const int lane_0 = opaque_zero();

  // pixel of step s on this scanline
  auto px = [&](int s) { return ROWS ? (sa > 0 ? s : W - 1 - s) : line; };
  auto py = [&](int s) { return ROWS ? line : (sa > 0 ? s : H - 1 - s); };

  // Uniform byte pointers of the next step to fetch, advanced by one step
  // per fetch; lanes add a constant 32-bit offset. A fetch past the
  // scanline's end loads nothing (its slot is never used).
  const int dxs = ROWS ? sa : 0, dys = ROWS ? 0 : sa;
  const int64_t vstep = ((int64_t)dys * sy + (int64_t)dxs * sx) * 4;
  const int64_t istep = ((int64_t)dys * W + dxs) * 4;
  const int64_t pstep = ((int64_t)dys * pitch + dxs) * 4;
  const int64_t v0 = ((int64_t)py(0) * sy + (int64_t)px(0) * sx) * 4;
  const char* fin = (const char*)in + v0;
  const char* facc = (const char*)out + v0;
  const char* fx1 = (const char*)x1p + ((int64_t)py(0) * pitch + kSgmPad + px(0) - kShift) * 4;
  const char* fx0 = (const char*)x0 + ((int64_t)py(0) * W + px(0)) * 4;
  char* pst = (char*)out + v0;  // the next step to store
  // the diagonal predecessor pixel's image values, relative to the pixel's
  const int64_t dgp = -((int64_t)dyd * pitch + dxd) * 4, dgi = -((int64_t)dyd * W + dxd) * 4;
  // Edge entries (first wavefront: the entry of step s >= 1 is the previous
  // block's (blk - 1) * S + s - 1; last wavefront: its own block's
  // blk * S + s): pointers of the next fetched, consumed and written step.
  const int64_t eprev = from_edge ? (int64_t)(blk - 1) * S : 0;
  const int64_t eown = to_edge ? (int64_t)blk * S : 0;
  const float* fe = sync.edge + eprev * NDP;
  const float* fem = sync.edge_mn + eprev;
  float* ce = sync.edge + eprev * NDP;
  float* cem = sync.edge_mn + eprev;
  float* te = sync.edge + eown * NDP;
  float* tem = sync.edge_mn + eown;
  const int lane_0x4 = 4 * lane_0;

  float rin[PF][DPL], rout[PF][DPL], rseg[PF][DPL], rx0[PF];
  // first wavefront: the previous block's values (edge buffer) and the
  // predecessor pixel's image values, PF - 1 steps ahead
  float er[PF][DPL], emn[PF], esg[PF][DPL], ex0[PF];

  // fetch of step s (called for s = 0, 1, 2, ... in order)
  auto fetch = [&](int slot, int s) {
    if (s >= S) return;
    if (act) {
      if (FULL) {
        bload(rsrc_at(fin), lane_v, rin[slot]);
        if (!FIRST) bload(rsrc_at(facc), lane_v, rout[slot]);
      } else {
        const auto ri = rsrc_at(fin), ra = rsrc_at(facc);
#pragma unroll
        for (int j = 0; j < DPL; j++) {
          rin[slot][j] = buffer_ldf(ri, 4 * joff[j], 0);
          if (!FIRST) rout[slot][j] = buffer_ldf(ra, 4 * joff[j], 0);
        }
      }
      bload(rsrc_at(fx1), lane_p, rseg[slot]);
      rx0[slot] = buffer_ldf(rsrc_at(fx0), lane_0x4, 0);
      if (from_edge && s >= 1) {
        const float* e = (const float*)((const char*)fe + lane_v);
#pragma unroll
        for (int j = 0; j < DPL; j++) er[slot][j] = __uint_as_float(ld_agent(e + j));
        emn[slot] = __uint_as_float(ld_agent(fem + lane_0));
        // the diagonal predecessor pixel, on the previous block's last scanline
        bload(rsrc_at(fx1 + dgp), lane_p, esg[slot]);
        ex0[slot] = buffer_ldf(rsrc_at(fx0 + dgi), lane_0x4, 0);
      }
    }
    fin += vstep;
    facc += vstep;
    fx1 += pstep;
    fx0 += istep;
    if (s >= 1) {
      fe += NDP;
      fem += 1;
    }
  };
  // edge entry of step s (the previous block's last scanline at step s - 1)
  auto edge_at = [&](int s) { return (int64_t)(blk - 1) * S + (s - 1); };

#pragma unroll
  for (int k = 0; k < PF - 1; k++) fetch(k, k);
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
  // The d-1 neighbour of lane 0's first disparity and the d+1 neighbour of
  // lane 63's last come out of the DPP shifts as 0 (bound_ctrl); their
  // penalties are +inf instead, which removes the term (d = -1 and
  // d = 64 DPL do not exist; 64 DPL > D - 1 is masked anyway).
  const bool lane_first = lane == 0, lane_last = lane == 63;
  const float inf = INFINITY;
  const float Mla = rl(ptab_a, 2), Mra = rl(ptab_a, 5), M2a = in_vgpr(rl(ptab_a, 8));
  const float Mld = rl(ptab_d, 2), Mrd = rl(ptab_d, 5), M2d = in_vgpr(rl(ptab_d, 8));
  const float Mla0 = lane_first ? inf : Mla, Mra1 = lane_last ? inf : Mra;
  const float Mld0 = lane_first ? inf : Mld, Mrd1 = lane_last ? inf : Mrd;
  const float Mla_v = in_vgpr(Mla), Mra_v = in_vgpr(Mra), Mld_v = in_vgpr(Mld),
              Mrd_v = in_vgpr(Mrd);
  const int taui = __float_as_int(tau_so + 0.f);
  const float fscale = (float)scale;
  float prev[DPL];
#pragma unroll
  for (int j = 0; j < DPL; j++) prev[j] = 0.f;
  float mn = 0.f;
  // after one wait has given up, later ones do not wait: a sweep stalls for
  // at most one spin_limit per wavefront
  bool gave_up = false;

  // One path's new values into cur, in float32: pv/mnp the predecessor's
  // values and minimum, (seg, pseg) the right-image pair, d1 the left-image
  // difference. A pair with a column outside the image reads the padding,
  // whose values (large, and different at every predecessor offset) make
  // D2 exceed tau as the reference's D2 = 10 does (post.cu:229-234).
  auto path_vals = [&](const float (&pv)[DPL], float mnp, const float (&cost)[DPL],
                       const float (&seg)[DPL], const float (&pseg)[DPL], float d1,
                       float ptab, float Ml, float Mr, float Ml0, float Mr1, float M2,
                       float (&cur)[DPL]) {
    // neighbours across lanes: wave shifts by one lane (0 where no source)
    const float left_nb = __int_as_float(__builtin_amdgcn_update_dpp(
        0, __float_as_int(pv[DPL - 1]), 0x138, 0xf, 0xf, true));
    const float right_nb = __int_as_float(__builtin_amdgcn_update_dpp(
        0, __float_as_int(pv[0]), 0x130, 0xf, 0xf, true));
    const int d1i = __builtin_amdgcn_readfirstlane(__float_as_int(d1));
    const bool lt1 = d1i < taui, gt1 = d1i > taui;
    const int cls = lt1 ? 0 : (gt1 ? 1 : 2);
    // D2's class test as one compare: sg D2 < tb with (sg, tb) = (1, tau)
    // when D1 < tau, (-1, -tau) when D1 > tau, (0, 0) (never true) otherwise
    const float sg = lt1 ? 1.f : (gt1 ? -1.f : 0.f);
    const float tb = lt1 ? tau_so : (gt1 ? -tau_so : 0.f);
    const float vAl = rl(ptab, cls), vAr = rl(ptab, 3 + cls), vA2 = rl(ptab, 6 + cls);
    const float vAl0 = lane_first ? inf : vAl, vAr1 = lane_last ? inf : vAr;
    const float qA = mnp + vA2, qM = mnp + M2;
#pragma unroll
    for (int j = 0; j < DPL; j++) {
      const float sj = DIR > 0 ? seg[j] : seg[DPL - 1 - j];
      const float pj = DIR > 0 ? pseg[j] : pseg[DPL - 1 - j];
      const bool cd = sg * fabsf(sj - pj) < tb;
      const float Pl = cd ? (j == 0 ? vAl0 : vAl) : (j == 0 ? Ml0 : Ml);
      const float Pr = cd ? (j == DPL - 1 ? vAr1 : vAr) : (j == DPL - 1 ? Mr1 : Mr);
      const float q = cd ? qA : qM;
      const float pl = j > 0 ? pv[j - 1] : left_nb;
      const float pr = j < DPL - 1 ? pv[j + 1] : right_nb;
      // a first pixel (no predecessor): pv = 0 and mnp = 0, so with
      // non-negative penalties c = 0 and the value is the cost itself
      const float c = fminf(fminf(pv[j], q), fminf(pl + Pl, pr + Pr));
      const float val = cost[j] + (c - mnp);
      cur[j] = (FULL || jok[j]) ? val : inf;
    }
  };
  // Minimum over the wavefront of non-negative values (path values are
  // costs in [0, 1] plus non-negative terms, or +inf): their bits order as
  // unsigned integers.
  // both paths' minima, their DPP steps interleaved so that each fills the
  // other's wait states
  auto vmin2 = [&](const float (&a)[DPL], const float (&b)[DPL], float& ma, float& mb) {
    uint32_t u = __float_as_uint(a[0]), w = __float_as_uint(b[0]);
#pragma unroll
    for (int j = 1; j < DPL; j++) {
      u = min(u, __float_as_uint(a[j]));
      w = min(w, __float_as_uint(b[j]));
    }
    u = min(u, (uint32_t)CBMV_DPP_U(u, 0xB1, 0xf));
    w = min(w, (uint32_t)CBMV_DPP_U(w, 0xB1, 0xf));
    u = min(u, (uint32_t)CBMV_DPP_U(u, 0x4E, 0xf));
    w = min(w, (uint32_t)CBMV_DPP_U(w, 0x4E, 0xf));
    u = min(u, (uint32_t)CBMV_DPP_U(u, 0x141, 0xf));
    w = min(w, (uint32_t)CBMV_DPP_U(w, 0x141, 0xf));
    u = min(u, (uint32_t)CBMV_DPP_U(u, 0x140, 0xf));
    w = min(w, (uint32_t)CBMV_DPP_U(w, 0x140, 0xf));
    u = min(u, (uint32_t)CBMV_DPP_U(u, 0x142, 0xa));
    w = min(w, (uint32_t)CBMV_DPP_U(w, 0x142, 0xa));
    u = min(u, (uint32_t)CBMV_DPP_U(u, 0x143, 0xc));
    w = min(w, (uint32_t)CBMV_DPP_U(w, 0x143, 0xc));
    ma = __uint_as_float(__builtin_amdgcn_readlane(u, 63));
    mb = __uint_as_float(__builtin_amdgcn_readlane(w, 63));
  };

  auto step = [&](int k, int s) {
    __builtin_amdgcn_sched_barrier(0);
    const int kp = (k + PF - 1) % PF;
    const int buf = s & 1;
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
      float* c = (float*)((char*)ce + lane_v);
#pragma unroll
      for (int j = 0; j < DPL; j++) c[j] = __uint_as_float(kEdgeEmpty);
      if (lane == 0) *cem = __uint_as_float(kEdgeEmpty);
      ce += NDP;
      cem += 1;
    } else {
#pragma unroll
      for (int j = 0; j < DPL; j++) {
        pd[j] = sh_d[buf ^ 1][wv - 1][dbase + j];
        psg[j] = sh_s[buf ^ 1][wv - 1][dbase + j];
      }
      mnd = sh_mn[buf ^ 1][wv - 1];
      px0 = sh_x0[buf ^ 1][wv - 1];
    }
    float ca[DPL], cdg[DPL];
    path_vals(prev, mn, rin[k], rseg[k], rseg[kp], fabsf(rx0[k] - rx0[kp]), ptab_a, Mla_v,
              Mra_v, Mla0, Mra1, M2a, ca);
    path_vals(pd, mnd, rin[k], rseg[k], psg, fabsf(rx0[k] - px0), ptab_d, Mld_v, Mrd_v, Mld0,
              Mrd1, M2d, cdg);
    float o[DPL];
#pragma unroll
    for (int j = 0; j < DPL; j++) {
      const float a = FIRST ? ca[j] : rout[k][j] + ca[j];
      o[j] = LAST ? (a + cdg[j]) * fscale : a + cdg[j];
      prev[j] = ca[j];
    }
    float mdo;
    vmin2(ca, cdg, mn, mdo);
    if (act) {
      const auto rs = rsrc_at(pst);
      if (FULL) {
        bstore(rs, lane_v, o);
      } else {
#pragma unroll
        for (int j = 0; j < DPL; j++)
          if (jok[j]) buffer_stf(o[j], rs, 4 * joff[j], 0);
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
        float* t = (float*)((char*)te + lane_v);
#pragma unroll
        for (int j = 0; j < DPL; j++) st_agent(t + j, cdg[j]);
        if (lane == 0) st_agent(tem + lane_0, mdo);
      }
    }
    pst += vstep;
    te += NDP;
    tem += 1;
    __syncthreads();
  };

  int s0 = 0;
  for (; s0 + PF <= S; s0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      step(k, s0 + k);
      fetch((k + PF - 1) % PF, s0 + k + PF - 1);
    }
  }
#pragma unroll
  for (int k = 0; k < PF; k++) {
    if (s0 + k < S) {
      step(k, s0 + k);
      fetch((k + PF - 1) % PF, s0 + k + PF - 1);
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
