// Persistent, software-pipelined stage kernels of the HEOM integrator for
// systems with N <= 32 (the matrix-core GEMM form).  Included by the
// per-method translation units.
//
// A block owns a sequence of tiles (T consecutive nodes each): tile
// blockIdx.x, then + gridDim.x, ...  The per-block constants (H as MFMA
// A-fragments, Delta and the edge-prefactor tables) are staged in LDS once.
// While tile i is combined, the global loads of tile i+1 are in flight:
//
//   top of iteration i:  elementwise part of tile i from registers -> LDS
//                        issue the epilogue loads of tile i and the input
//                        and edge-record loads of tile i+1
//   barrier              -(i/hbar) H Y of tile i on the matrix cores (LDS)
//   barrier              issue the neighbour gathers of tile i+1 (their edge
//                        records arrived during the GEMM)
//                        epilogue of tile i: k = W + W^dagger, stage update
//
// LDS holds Y and W of two tiles (alternating), so no barrier is needed
// between the epilogue of tile i and the elementwise part of tile i+1.
//
// Elements are numbered e = t*N*N + b*N + a (node t of the tile, column b,
// row a), thread tid owns e = tid + i*blockDim.x (i < EPT), so every global
// access of a wavefront is contiguous within a column.  In LDS, Y and W use
// column stride N + 1.
//
// Edge records (built on the host): edge[(n*N + b)*MPS + q] = {id of
// n + e_j, id of n - e_j, j*L + n_j, 0} for the q-th mode j coupled to site
// b, ids -1 where the neighbour does not exist, and {-1, -1, -1, 0} for
// unused slots.  The RHS it evaluates is the same as rhs_phase_mfma's:
//   W(a,b) = -(i/hbar) (H Y)(a,b) - (mu/2 + (Delta_a + Delta_b)/2 [a != b]) Y(a,b)
//            + sum_q i s_q Y_{n+e_j}(a,b) - col_q Y_{n-e_j}(a,b),
//   k = W + W^dagger.
#ifndef PHI_AMD_HEOM_STREAM_H_
#define PHI_AMD_HEOM_STREAM_H_

#include "heom_device.h"

namespace phi {
namespace dev {

// H fragments in LDS: hf[(mt*KS + st)*64 + lane] = H(16 mt + lane%16,
// (lane/16)*KS + st) (zero outside N), real parts then imaginary parts.
__host__ __device__ inline int StreamFragCount(int N) {
  return ((N + 15) >> 4) * ((N + 3) >> 2) * 64;
}

__host__ __device__ inline size_t StreamLdsBytes(int N, int D, int L, int T, size_t real) {
  const size_t V = 2 * real, NP = (size_t)N * (N + 1);
  size_t b = V * (4 * NP * T + (size_t)D * L);        // Y[2], W[2], pcol
  b += real * ((size_t)D * L + N + 2 * (size_t)StreamFragCount(N));  // ns, delta, hf
  return Align16(b);
}

template <typename TB>
struct StreamLds {
  typedef typename C2<TB>::type V;
  V* Y;       // two tiles, the second at Y + bstride
  V* W;       // two tiles, the second at W + bstride
  int bstride;
  V* pcol;
  TB* ns;
  TB* delta;
  TB* hfr;
  TB* hfi;
};

template <typename TB>
__device__ __forceinline__ StreamLds<TB> stream_carve(unsigned char* smem, int N, int D, int L,
                                                     int T) {
  typedef typename C2<TB>::type V;
  const int NP = N * (N + 1);
  StreamLds<TB> s;
  V* p = reinterpret_cast<V*>(smem);
  s.bstride = NP * T;
  s.Y = p;
  s.W = p + 2 * NP * T;
  s.pcol = p + 4 * NP * T;
  s.ns = reinterpret_cast<TB*>(s.pcol + D * L);
  s.delta = s.ns + D * L;
  s.hfr = s.delta + N;
  s.hfi = s.hfr + StreamFragCount(N);
  return s;
}

// -(i/hbar) H Y of one tile on the matrix cores, added into W.  Lane l of a
// 16x16x4 MFMA holds A[l%16][l/16] and B[l/16][l%16]; the K dimension is
// split so that lane group g = l/16 covers rows g*KS .. g*KS + KS - 1.
// Work items are (16-column slice, 16-row tile of H) pairs dealt to the
// waves round robin.  A wavefront's first two items have the same place in
// every tile: their column, tile row and LDS offset are computed once per
// kernel (GemmItems) and a tile only checks its column count.
struct GemmItems {
  int j[2];     // tile column of this lane
  int mt[2];    // 16-row tile of H
  int off[2];   // LDS offset of the column (Y and W share the layout)
};

__device__ __forceinline__ GemmItems stream_gemm_items(const RhsArgs& A) {
  GemmItems g;
  const int N = A.N, NP = N * (N + 1), MT = (N + 15) >> 4;
  const int c16 = threadIdx.x & 15, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int item = wave + i * nwaves;
    const int slice = item / MT;
    g.mt[i] = item - slice * MT;
    g.j[i] = slice * 16 + c16;
    const int tj = fdiv(g.j[i], A.mN);
    g.off[i] = tj * NP + (g.j[i] - tj * N) * (N + 1);
  }
  return g;
}

template <typename TB>
__device__ __forceinline__ void stream_gemm_item(const RhsArgs& A, const StreamLds<TB>& s,
                                                 const typename C2<TB>::type* ycol,
                                                 typename C2<TB>::type* wcol, bool jok, int mt) {
  typedef typename C2<TB>::type V;
  typedef Mfma<TB> M;
  const int N = A.N;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int KS = (N + 3) >> 2;
  const TB ih = (TB)A.inv_hbar;
  const bool realh = A.real_h != 0;
  typename M::acc xr = M::zero(), xi = M::zero();
  for (int st = 0; st < KS; ++st) {
    const int k = g * KS + st;
    V y = mk<TB>(0, 0);
    if (jok && k < N) y = ycol[k];
    const TB hr = s.hfr[(mt * KS + st) * 64 + lane];
    xr = M::mac(hr, y.x, xr);
    xi = M::mac(hr, y.y, xi);
    if (!realh) {
      const TB hi = s.hfi[(mt * KS + st) * 64 + lane];
      xr = M::mac(-hi, y.y, xr);
      xi = M::mac(hi, y.x, xi);
    }
  }
  if (jok) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * mt + M::row(g, r);
      if (row < N) {
        V w = wcol[row];
        w.x += ih * xi[r];
        w.y -= ih * xr[r];
        wcol[row] = w;
      }
    }
  }
}

template <typename TB>
__device__ __forceinline__ void stream_gemm(const RhsArgs& A, const StreamLds<TB>& s,
                                            const typename C2<TB>::type* Y,
                                            typename C2<TB>::type* W, int nt,
                                            const GemmItems& gi) {
  const int N = A.N;
  const int NP = N * (N + 1);
  const int c16 = threadIdx.x & 15;
  const int MT = (N + 15) >> 4;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int ncols = nt * N;
  const int nitems = ((ncols + 15) >> 4) * MT;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (wave + i * nwaves < nitems)
      stream_gemm_item<TB>(A, s, Y + gi.off[i], W + gi.off[i], gi.j[i] < ncols, gi.mt[i]);
  }
  // Further items (shapes with more than two per wavefront).
  for (int item = wave + 2 * nwaves; item < nitems; item += nwaves) {
    const int slice = item / MT;
    const int mt = item - slice * MT;
    const int j = slice * 16 + c16;
    const bool jok = j < ncols;
    const int tj = jok ? fdiv(j, A.mN) : 0;
    const int off = tj * NP + (j - tj * N) * (N + 1);
    stream_gemm_item<TB>(A, s, Y + off, W + off, jok, mt);
  }
}

// Registers of the next tile in flight (input element, mu, edge records and
// neighbour gathers of each owned element).  mu stays as loaded until the
// tile is combined: converting it at load time would make the wavefront wait
// for the load before the GEMM, which is what the prefetch is meant to hide.
template <typename TB, typename TI, int EPT, int MPS>
struct StreamRegs {
  typedef typename C2<TI>::type VI;
  VI y[EPT];
  double mu[EPT];
  int4 rec[EPT][MPS];
  VI gn[EPT][MPS];
  VI gp[EPT][MPS];
};

// Element e of this thread: node te, in-node offset r = b*N + a.
__device__ __forceinline__ void stream_coords(const RhsArgs& A, int i, int* te, int* r, int* b,
                                              int* a) {
  const int e = threadIdx.x + i * blockDim.x;
  *te = fdiv(e, A.mNE);
  *r = e - *te * A.NE;
  *b = fdiv(*r, A.mN);
  *a = *r - *b * A.N;
}

// The coordinates of a thread's elements, the same in every tile (computed
// once per kernel).
template <int EPT>
struct StreamCoords {
  int te[EPT], r[EPT], b[EPT], a[EPT];
};

template <int EPT>
__device__ __forceinline__ StreamCoords<EPT> stream_coords_all(const RhsArgs& A) {
  StreamCoords<EPT> c;
#pragma unroll
  for (int i = 0; i < EPT; ++i) stream_coords(A, i, &c.te[i], &c.r[i], &c.b[i], &c.a[i]);
  return c;
}

// A tile: slab node `base` holds local node `local` of its member (the
// member's nodes are slab nodes base - local .. base - local + count - 1),
// nt nodes.  A single hierarchy is one member with count = num_nodes.
struct TileInfo {
  int base, local, nt, member, count;
};

// ENS: the tile comes from the descriptor list (an ensemble).  The tile index
// is uniform over the block, so the descriptor is read through the scalar
// cache (a constant-address-space load): it does not enter the in-order
// vector-memory counter, so waiting for it never waits for the vector loads
// in flight.
template <bool ENS>
__device__ __forceinline__ TileInfo stream_tile(const RhsArgs& A, int tile) {
  TileInfo t;
  if constexpr (ENS) {
    typedef const __attribute__((address_space(4))) int* ConstInt;
    const ConstInt d = (ConstInt)(A.tiles + tile);
    t.base = d[0];
    t.local = d[1];
    t.nt = d[2];
    t.member = d[3];
    t.count = d[4];
  } else {
    t.base = tile * A.tile;
    t.local = t.base;
    t.nt = min(A.tile, A.num_nodes - t.base);
    t.member = 0;
    t.count = A.num_nodes;
  }
  return t;
}

// Loads are unconditional: an element beyond the tile (a short last tile,
// or a thread past the tile's elements) loads those of the tile's last node
// instead, and an absent neighbour's gather loads the node's own element;
// the values are never used.  Divergent branches around loads would make
// the compiler's wait counting assume the worst at every join and drain all
// loads in flight, prefetches included.

// Loads of a tile that need no other load first: its input elements, mu
// and the edge records (tables indexed by the member-local node id).
template <typename TB, typename TI, int EPT, int MPS>
__device__ __forceinline__ void stream_load1(const RhsArgs& A, const TI* __restrict__ in,
                                             const TileInfo& t, const StreamCoords<EPT>& C,
                                             StreamRegs<TB, TI, EPT, MPS>& R) {
  typedef typename C2<TI>::type VI;
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int te = min(C.te[i], t.nt - 1);
    const int ln = t.local + te;
    R.y[i] = reinterpret_cast<const VI*>(in)[(size_t)(t.base + te) * A.NE + C.r[i]];
    R.mu[i] = A.mu[ln];
    const int4* ed = A.edge + ((size_t)ln * A.N + C.b[i]) * MPS;
#pragma unroll
    for (int q = 0; q < MPS; ++q) R.rec[i][q] = ed[q];
  }
}

// Neighbour gathers of a tile (after its edge records arrived).  In an
// ensemble, local ids become slab ids and neighbours beyond the member's
// depth do not exist; a single hierarchy's ids are slab ids already.
template <typename TB, typename TI, int EPT, int MPS, bool ENS>
__device__ __forceinline__ void stream_load2(const RhsArgs& A, const TI* __restrict__ in,
                                             const TileInfo& t, const StreamCoords<EPT>& C,
                                             StreamRegs<TB, TI, EPT, MPS>& R) {
  typedef typename C2<TI>::type VI;
  const VI* v = reinterpret_cast<const VI*>(in);
  const int off = t.base - t.local;
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int self = t.base + min(C.te[i], t.nt - 1);
#pragma unroll
    for (int q = 0; q < MPS; ++q) {
      int4 c = R.rec[i][q];
      if constexpr (ENS) {
        c.x = (c.x >= 0 && c.x < t.count) ? c.x + off : -1;
        c.y = (c.y >= 0 && c.y < t.count) ? c.y + off : -1;
        R.rec[i][q] = c;
      }
      const int jn = c.x >= 0 ? c.x : self;
      const int jp = c.y >= 0 ? c.y : self;
      R.gn[i][q] = v[(size_t)jn * A.NE + C.r[i]];
      R.gp[i][q] = v[(size_t)jp * A.NE + C.r[i]];
    }
  }
}

// Element f of the MFMA A-fragments of H (zero padding outside N x N).
__device__ __forceinline__ double2 stream_h_frag(const RhsArgs& A, const double2* H, int f) {
  const int N = A.N;
  const int KS = (N + 3) >> 2;
  const int lane = f & 63, ms = f >> 6;
  const int mt = ms / KS, st = ms - mt * KS;
  const int row = 16 * mt + (lane & 15), k = (lane >> 4) * KS + st;
  const bool ok = row < N && k < N;
  const double2 h = H[ok ? row + k * N : 0];
  return ok ? h : make_double2(0., 0.);
}

// H of `member` (shifted by its mean diagonal) as MFMA A-fragments in LDS.
template <typename TB>
__device__ __forceinline__ void stream_load_h(const RhsArgs& A, const StreamLds<TB>& s,
                                              int member) {
  const int nf = StreamFragCount(A.N);
  const double2* H = A.H + (size_t)member * A.NE;
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    const double2 h = stream_h_frag(A, H, f);
    s.hfr[f] = (TB)h.x;
    s.hfi[f] = (TB)h.y;
  }
}

// One element per thread allows blocks of up to 1024 threads (a whole
// 24 x 24 node per block); more elements per thread keep blocks at <= 256.
template <typename TB, typename TI, int EPT, int MPS, bool ENS, class EPI>
__global__ void __launch_bounds__(EPT == 1 ? 1024 : 256)
    stream_stage(RhsArgs A, const TI* __restrict__ in, EPI epi) {
  typedef typename C2<TB>::type V;
  typedef StreamRegs<TB, TI, EPT, MPS> Regs;
  typedef typename EPI::Pre Pre;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, NE = A.NE, D = A.D, L = A.L, T = A.tile;
  const int NP = N * (N + 1);
  const StreamLds<TB> s = stream_carve<TB>(smem, N, D, L, T);
  // Tiles of this block.  Workgroups are dealt to the 8 XCDs round robin
  // (block b on XCD b % 8), each XCD with its own L2.
  //  * Ensemble: the descriptor list is cut into X = min(G, 8) contiguous
  //    runs, run x for the blocks b with b % X == x, which take every G_x-th
  //    tile of it (G_x blocks in the run).  The blocks of one XCD therefore
  //    work together on a window of G_x consecutive tiles, inside one or two
  //    members, and every neighbour gather of that window (a member's
  //    neighbours lie anywhere in its tier-major id range) is served by that
  //    XCD's L2, which holds the member's stage input.
  //  * Single hierarchy: every G-th tile of the hierarchy, starting from an
  //    XCD-aware slot, so the slots of XCD x form one contiguous run
  //    [start_x, start_x + count_x) and the neighbour ADOs that nearby tiles
  //    share are fetched into one XCD's L2 instead of several.
  const int ntiles = ENS ? A.ntiles : (A.num_nodes + T - 1) / T;
  int tile, tile_end, step;
  if constexpr (ENS) {
    const int G = gridDim.x, X = G < 8 ? G : 8;
    const int x = blockIdx.x % X;
    tile = (int)(((long long)x * ntiles) / X) + (int)(blockIdx.x / X);
    tile_end = (int)(((long long)(x + 1) * ntiles) / X);
    step = G / X + (x < G % X ? 1 : 0);
  } else {
    const int G = gridDim.x, x = blockIdx.x & 7, base = G >> 3, r = G & 7;
    tile = x * base + min(x, r) + (int)(blockIdx.x >> 3);
    tile_end = ntiles;
    step = G;
  }
  if (tile >= tile_end) return;

  // Prologue: every load that needs no other load is issued first (the
  // first tile's elements, mu and edge records; each thread's first share
  // of the per-block constants and of H's fragments), then the LDS stores,
  // then the neighbour gathers.  Two trips to memory before the first
  // barrier instead of one per table: short kernels are mostly prologue.
  double emax = 0;
  const StreamCoords<EPT> C = stream_coords_all<EPT>(A);
  Regs cur;
  TileInfo ti = stream_tile<ENS>(A, tile);
  int hmember = ti.member;
  stream_load1<TB, TI, EPT, MPS>(A, in, ti, C, cur);
  const int DL = D * L, nf = StreamFragCount(N);
  const double2* Hm = A.H + (size_t)hmember * NE;
  const int e0 = threadIdx.x;
  V pc0 = mk<TB>((TB)0, (TB)0);
  double ns0 = 0, dl0 = 0;
  double2 h0 = make_double2(0., 0.);
  if (e0 < DL) {
    pc0 = ld<TB, double>(reinterpret_cast<const double*>(A.prev_col), e0);
    ns0 = A.next_s[e0];
  }
  if (e0 < N) dl0 = A.delta[e0];
  if (e0 < nf) h0 = stream_h_frag(A, Hm, e0);
  if (e0 < DL) {
    s.pcol[e0] = pc0;
    s.ns[e0] = (TB)ns0;
  }
  if (e0 < N) s.delta[e0] = (TB)dl0;
  if (e0 < nf) {
    s.hfr[e0] = (TB)h0.x;
    s.hfi[e0] = (TB)h0.y;
  }
  for (int e = e0 + (int)blockDim.x; e < DL; e += blockDim.x) {
    s.pcol[e] = ld<TB, double>(reinterpret_cast<const double*>(A.prev_col), e);
    s.ns[e] = (TB)A.next_s[e];
  }
  for (int e = e0 + (int)blockDim.x; e < N; e += blockDim.x) s.delta[e] = (TB)A.delta[e];
  for (int f = e0 + (int)blockDim.x; f < nf; f += blockDim.x) {
    const double2 h = stream_h_frag(A, Hm, f);
    s.hfr[f] = (TB)h.x;
    s.hfi[f] = (TB)h.y;
  }
  stream_load2<TB, TI, EPT, MPS, ENS>(A, in, ti, C, cur);
  // Every prologue load has arrived before the loop (the table loops above
  // have data-dependent trip counts; without this wait the compiler's load
  // counting could not tell, inside the loop, which loads are still in
  // flight, and would drain them all before each use).
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();  // per-block tables visible
  // Per-element constants of every tile: the Matsubara correction
  // (Delta_a + Delta_b)/2 off the diagonal and the LDS offsets of the
  // element and of its transpose.
  TB fd[EPT];
  int oo[EPT], ot[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int te = min(C.te[i], T - 1), b = C.b[i], a = C.a[i];
    fd[i] = a != b ? (TB)0.5 * (s.delta[a] + s.delta[b]) : (TB)0;
    oo[i] = te * NP + b * (N + 1) + a;
    ot[i] = te * NP + a * (N + 1) + b;
  }
  double2 hp[2] = {make_double2(0., 0.), make_double2(0., 0.)};
  const GemmItems gi = stream_gemm_items(A);

  for (int buf = 0; tile < tile_end; tile += step, buf ^= 1) {
    const int nt = ti.nt, base = ti.base;
    const int tot = nt * NE;
    V* Yb = s.Y + buf * s.bstride;
    V* Wb = s.W + buf * s.bstride;
    // H of a new member (the previous tile's product finished before the
    // last barrier; the next barrier publishes the new fragments), from the
    // registers the previous iteration loaded (ensemble launches have at
    // least half as many threads as fragment elements).
    if (ENS && ti.member != hmember) {
      hmember = ti.member;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int f = threadIdx.x + j * blockDim.x;
        if (f < nf) {
          s.hfr[f] = (TB)hp[j].x;
          s.hfi[f] = (TB)hp[j].y;
        }
      }
    }
    // (1) epilogue operands of every owned element (clamped), then the
    // elementwise part of this tile -> LDS
    Pre pre[EPT];
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const size_t g = (size_t)(base + min(C.te[i], nt - 1)) * NE + C.r[i];
      if (EPI::kInIsState) pre[i] = epi.from_state(make_double2((double)cur.y[i].x,
                                                                (double)cur.y[i].y));
      else pre[i] = epi.pre(g);
    }
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const V y = mk<TB>((TB)cur.y[i].x, (TB)cur.y[i].y);
      const TB f = (TB)(0.5 * cur.mu[i]) + fd[i];
      TB ax = -f * y.x, ay = -f * y.y;
#pragma unroll
      for (int q = 0; q < MPS; ++q) {
        // absent neighbours (id -1) contribute nothing; their table index
        // is -1 too, read as entry 0 and multiplied by 0
        const int4 c = cur.rec[i][q];
        const int z = max(c.z, 0);
        const TB sc = c.x >= 0 ? s.ns[z] : (TB)0;
        V pc = s.pcol[z];
        if (c.y < 0) pc = mk<TB>((TB)0, (TB)0);
        const TB gx = (TB)cur.gn[i][q].x, gy = (TB)cur.gn[i][q].y;
        const TB hx = (TB)cur.gp[i][q].x, hy = (TB)cur.gp[i][q].y;
        ax -= sc * gy + (pc.x * hx - pc.y * hy);
        ay += sc * gx - (pc.x * hy + pc.y * hx);
      }
      if (threadIdx.x + i * blockDim.x < tot) {
        Yb[oo[i]] = y;
        Wb[oo[i]] = mk<TB>(ax, ay);
      }
    }
    // (2) first loads of the next tile (the current one again after the
    // last, loaded and never combined, so that no load depends on a branch)
    const int ntile = tile + step;
    const bool more = ntile < tile_end;
    const TileInfo nti = stream_tile<ENS>(A, more ? ntile : tile);
    stream_load1<TB, TI, EPT, MPS>(A, in, nti, C, cur);
    if constexpr (ENS) {
      // The next tile's H fragments (L2 hits, in flight with the tile's
      // other loads), used when its member differs.
      const double2* Hn = A.H + (size_t)nti.member * NE;
#pragma unroll
      for (int j = 0; j < 2; ++j)
        hp[j] = stream_h_frag(A, Hn, min((int)threadIdx.x + j * (int)blockDim.x, nf - 1));
    }
    __syncthreads();
    // (3) matrix-core product
    stream_gemm<TB>(A, s, Yb, Wb, nt, gi);
    __syncthreads();
    // (4) neighbour gathers of the next tile (its edge records arrived
    // during the product)
    stream_load2<TB, TI, EPT, MPS, ENS>(A, in, nti, C, cur);
    // (5) epilogue of this tile
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      if (threadIdx.x + i * blockDim.x < tot) {
        const V p = Wb[oo[i]];
        const V q = Wb[ot[i]];
        double2 k;
        k.x = (double)(p.x + q.x);
        k.y = (double)(p.y - q.y);
        epi.apply((size_t)(base + C.te[i]) * NE + C.r[i], k, pre[i], &emax);
      }
    }
    ti = nti;
  }
  // Epilogues that reduce (the RKF45 error norm): max over the block, one
  // atomic per block.
  if constexpr (EPI::kReduce) {
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) {
      const double v = __shfl_xor(emax, o, 64);
      emax = v > emax ? v : emax;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = emax;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
      epi.finalize(m);
    }
  }
}

// --- RK4 epilogues ---------------------------------------------------------------
// Nested form: stage s < 3 writes v = rho + c k (c = dt/(4-s)), stage 3
// writes rho += dt k.  Stage 0 reads rho as its input (kInIsState).
template <typename TB, int STAGE, bool IN_STATE>
struct EpiNested {
  struct Pre { double2 x; };
  static constexpr bool kInIsState = IN_STATE;
  static constexpr bool kReduce = false;
  double* rho;
  TB* yout;
  double c;
  __device__ __forceinline__ Pre pre(size_t g) const { return Pre{ld2_nt(rho, g)}; }
  __device__ __forceinline__ Pre from_state(double2 x) const { return Pre{x}; }
  __device__ __forceinline__ void apply(size_t g, double2 k, const Pre& p, double*) const {
    if (STAGE < 3) st_nt<TB>(yout, g, p.x.x + c * k.x, p.x.y + c * k.y);
    else st_nt<double>(rho, g, p.x.x + c * k.x, p.x.y + c * k.y);
  }
  __device__ __forceinline__ void finalize(double) const {}
};

// Increment form of the same step: with K = L rho, the stage inputs are
// increments instead of rho + increment,
//   s0: K = L rho                       s1: D = dt/3 (K + dt/4 L K)
//   s2: D' = dt/2 (K + L D)             s3: rho += dt (K + L D'),
// which is the nested step (Y1 = rho + dt/4 K, Y2 = rho + D, Y3 = rho + D')
// rearranged.  The FP64 state is read by the first and last stage only
// (N^2 (48 + 8 s) bytes per node-step against N^2 (80 + 6 s): fewer for
// FP32 bulk, the same for FP64), and the bulk copies hold increments, which
// round far less than rho + increment.  K, D and D' are Hermitian when rho
// is (linear combinations of L applied to Hermitian ADOs).  Stages 0 and 1
// take their own input element as the epilogue operand (kInIsState).
template <typename TB, int STAGE>
struct EpiIncr {
  typedef typename C2<TB>::type V;
  struct Pre { double2 x; V q; };
  static constexpr bool kInIsState = STAGE <= 1;
  static constexpr bool kReduce = false;
  double* rho;
  const TB* kk;   // K (stages 2 and 3)
  TB* yout;
  double dt;
  __device__ __forceinline__ Pre pre(size_t g) const {
    Pre p;
    p.q = ldv_nt<TB>(kk, g);
    if (STAGE == 3) p.x = ld2_nt(rho, g);
    return p;
  }
  __device__ __forceinline__ Pre from_state(double2 x) const {
    Pre p;
    p.x = x;
    return p;
  }
  // Stages 0-2 write bulk values and compute in the bulk precision; the
  // state update (stage 3) is FP64.
  __device__ __forceinline__ void apply(size_t g, double2 k, const Pre& p, double*) const {
    const TB kx = (TB)k.x, ky = (TB)k.y;
    if (STAGE == 0) {
      stv_nt<TB>(yout, g, mk<TB>(kx, ky));
    } else if (STAGE == 1) {
      const TB c1 = (TB)(dt / 3), c0 = (TB)(dt / 4);
      const TB x = (TB)p.x.x, y = (TB)p.x.y;
      stv_nt<TB>(yout, g, mk<TB>(c1 * (x + c0 * kx), c1 * (y + c0 * ky)));
    } else if (STAGE == 2) {
      const TB c2 = (TB)(dt / 2);
      stv_nt<TB>(yout, g, mk<TB>(c2 * (p.q.x + kx), c2 * (p.q.y + ky)));
    } else {
      st_nt<double>(rho, g, p.x.x + dt * ((double)p.q.x + k.x),
                    p.x.y + dt * ((double)p.q.y + k.y));
    }
  }
  __device__ __forceinline__ void finalize(double) const {}
};

// Classic form (PHI's order):
//   s0: acc = rho + dt/6 k, v = rho + dt/2 k     s1: acc += dt/3 k, v = rho + dt/2 k
//   s2: acc += dt/3 k, v = rho + dt k            s3: rho = acc + dt/6 k
template <typename TB, int STAGE, bool IN_STATE>
struct EpiClassic {
  struct Pre { double2 x, q; };
  static constexpr bool kInIsState = IN_STATE;
  static constexpr bool kReduce = false;
  double* rho;
  double* acc;
  TB* yout;
  double dt;
  __device__ __forceinline__ void finalize(double) const {}
  __device__ __forceinline__ Pre pre(size_t g) const {
    Pre p;
    if (STAGE <= 2) p.x = ld2(rho, g);
    if (STAGE >= 1) p.q = ld2(acc, g);
    return p;
  }
  __device__ __forceinline__ Pre from_state(double2 x) const {
    Pre p;
    p.x = x;
    return p;
  }
  __device__ __forceinline__ void apply(size_t g, double2 k, const Pre& p, double*) const {
    if (STAGE == 0) {
      st<double>(acc, g, p.x.x + (dt / 6) * k.x, p.x.y + (dt / 6) * k.y);
      st<TB>(yout, g, p.x.x + (dt / 2) * k.x, p.x.y + (dt / 2) * k.y);
    } else if (STAGE <= 2) {
      st<double>(acc, g, p.q.x + (dt / 3) * k.x, p.q.y + (dt / 3) * k.y);
      const double c = STAGE == 1 ? dt / 2 : dt;
      st<TB>(yout, g, p.x.x + c * k.x, p.x.y + c * k.y);
    } else {
      st<double>(rho, g, p.q.x + (dt / 6) * k.x, p.q.y + (dt / 6) * k.y);
    }
  }
};

// Classic order with the accumulator and a copy of rho in the bulk
// precision (kRk4ClassicBulk, the learned closure's FP32 runs): the stage
// inputs are bulk values in PHI's classic form too, so forming them from
// the bulk copy changes only their rounding, and the accumulator holds only
// the step's increment; rho is read by the first and last stage only:
//   s0: acc = dt/6 k, rb = rho, v = rho + dt/2 k     s1: acc += dt/3 k, v = rb + dt/2 k
//   s2: acc += dt/3 k, v = rb + dt k                  s3: rho += acc + dt/6 k
// Per element and step 2 x 16 + 14 x s bytes (s = bulk complex size) against
// the classic form's 10 x 16 + 6 x s.
template <typename TB, int STAGE, bool IN_STATE>
struct EpiClassicBulk {
  typedef typename C2<TB>::type V;
  struct Pre { double2 x; V q, r; };
  static constexpr bool kInIsState = IN_STATE;
  static constexpr bool kReduce = false;
  double* rho;
  TB* acc;
  TB* rb;
  TB* yout;
  double dt;
  __device__ __forceinline__ void finalize(double) const {}
  __device__ __forceinline__ Pre pre(size_t g) const {
    Pre p;
    if (STAGE == 0 || STAGE == 3) p.x = ld2_nt(rho, g);
    if (STAGE == 1 || STAGE == 2) p.r = ldv_nt<TB>(rb, g);
    if (STAGE >= 1) p.q = ldv_nt<TB>(acc, g);
    return p;
  }
  __device__ __forceinline__ Pre from_state(double2 x) const {
    Pre p;
    p.x = x;
    return p;
  }
  __device__ __forceinline__ void apply(size_t g, double2 k, const Pre& p, double*) const {
    const TB kx = (TB)k.x, ky = (TB)k.y;
    if (STAGE == 0) {
      const TB c6 = (TB)(dt / 6);
      stv_nt<TB>(acc, g, mk<TB>(c6 * kx, c6 * ky));
      stv_nt<TB>(rb, g, mk<TB>((TB)p.x.x, (TB)p.x.y));
      st_nt<TB>(yout, g, p.x.x + (dt / 2) * k.x, p.x.y + (dt / 2) * k.y);
    } else if (STAGE <= 2) {
      const TB c3 = (TB)(dt / 3), c = (TB)(STAGE == 1 ? dt / 2 : dt);
      stv_nt<TB>(acc, g, mk<TB>(p.q.x + c3 * kx, p.q.y + c3 * ky));
      stv_nt<TB>(yout, g, mk<TB>(p.r.x + c * kx, p.r.y + c * ky));
    } else {
      st_nt<double>(rho, g, p.x.x + ((double)p.q.x + (dt / 6) * k.x),
                    p.x.y + ((double)p.q.y + (dt / 6) * k.y));
    }
  }
};

// --- RKF45 epilogue ------------------------------------------------------------
// Stage s < 5 stores k_s (bulk precision) and the next stage input
//   tmp = rho + sum_{j <= s} b[s][j] k_j
// (k_s as stored, so later stages see the same value; left-to-right sums as
// PHI, hierarchy_integrator.cc:2019-2253).  Stage 5 forms
//   y4 = rho + c4 . (k0, k2, k3, k4),  y5 = rho + c5 . (k0, k2, k3, k4, k5),
// stores y5 (FP64) and reduces max |y4 - y5| (hierarchy_integrator.cc:
// 2317-2328) into *err (one atomic max per block on the bits of the
// non-negative double).
template <typename TB, int STAGE, bool IN_STATE>
struct EpiRkf45 {
  typedef typename C2<TB>::type V;
  struct Pre {
    double2 x;
    V k[5];
  };
  static constexpr bool kInIsState = IN_STATE;
  static constexpr bool kReduce = STAGE == 5;
  Rkf45Bufs<TB> B;
  const double* rho;
  Rkf45Coefs cf;
  __device__ __forceinline__ Pre pre(size_t g) const {
    Pre p;
    p.x = ld2(rho, g);
#pragma unroll
    for (int j = 0; j < 5; ++j)
      if (j < STAGE && !(STAGE == 5 && j == 1)) p.k[j] = ld<TB, TB>(B.k[j], g);
    return p;
  }
  __device__ __forceinline__ Pre from_state(double2 x) const {
    Pre p;
    p.x = x;
    return p;
  }
  __device__ __forceinline__ void apply(size_t g, double2 k, const Pre& p, double* emax) const {
    const double ksx = (double)(TB)k.x, ksy = (double)(TB)k.y;
    if (STAGE < 5) {
      st<TB>(B.k[STAGE], g, k.x, k.y);
      double sx = p.x.x, sy = p.x.y;
#pragma unroll
      for (int j = 0; j < 5; ++j)
        if (j < STAGE) {
          sx += (double)p.k[j].x * cf.b[STAGE][j];
          sy += (double)p.k[j].y * cf.b[STAGE][j];
        }
      sx += ksx * cf.b[STAGE][STAGE];
      sy += ksy * cf.b[STAGE][STAGE];
      st<TB>(B.tmp_out, g, sx, sy);
    } else {
      const V k0 = p.k[0], k2 = p.k[2], k3 = p.k[3], k4 = p.k[4];
      const double y4x = p.x.x + (double)k0.x * cf.c4[0] + (double)k2.x * cf.c4[1] +
                         (double)k3.x * cf.c4[2] + (double)k4.x * cf.c4[3];
      const double y4y = p.x.y + (double)k0.y * cf.c4[0] + (double)k2.y * cf.c4[1] +
                         (double)k3.y * cf.c4[2] + (double)k4.y * cf.c4[3];
      const double y5x = p.x.x + (double)k0.x * cf.c5[0] + (double)k2.x * cf.c5[1] +
                         (double)k3.x * cf.c5[2] + (double)k4.x * cf.c5[3] + ksx * cf.c5[4];
      const double y5y = p.x.y + (double)k0.y * cf.c5[0] + (double)k2.y * cf.c5[1] +
                         (double)k3.y * cf.c5[2] + (double)k4.y * cf.c5[3] + ksy * cf.c5[4];
      st<double>(B.rho_new, g, y5x, y5y);
      const double d = hypot(y4x - y5x, y4y - y5y);
      *emax = d > *emax ? d : *emax;
    }
  }
  __device__ __forceinline__ void finalize(double m) const { atomic_max_nonneg(B.err, m); }
};

}  // namespace dev
}  // namespace phi

#endif  // PHI_AMD_HEOM_STREAM_H_
