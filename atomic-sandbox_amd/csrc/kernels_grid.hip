// gfx950 kernels on the dense n^3 grid: the block stencil H.X, reductions, the tall-skinny Gram S^T U, block
// linear combinations, density accumulation, LSDA, nuclear potential, orbital guess, DIIS mix and HF forces.
// All kernels are written for 64-wide wavefronts; reductions are deterministic (fixed partial counts, fixed order).
#include "lira_internal.hpp"
#include "lsda.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <vector>

namespace lira {

void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t b) {
    release();
    if (b == 0) b = 256;
    LIRA_HIP(hipMalloc(&ptr, b));
    bytes = b;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// 8th-order second-derivative coefficients c1..c4 (qm/src/stencil.rs:24; c0 = -205/72 is folded into the
// differences: sum_m c_m[(f+ - f) + (f- - f)] per axis equals c0 f + sum_m c_m (f+ + f-) since c0 = -2 sum c_m).
// Kernels use them as T(Cm): an f32 stencil computes in f32 throughout (a bare double constant would promote every
// tap product to f64).
#define C1 (8.0 / 5.0)
#define C2 (-1.0 / 5.0)
#define C3 (8.0 / 315.0)
#define C4 (-1.0 / 560.0)

// ---- buffer addressing of the stencil fields ----
// The stencils read and write whole fields through raw buffer descriptors whose range is one field. A load at a byte
// offset outside [0, range) returns 0 -- the Dirichlet zero of every out-of-box neighbour -- and a store there is
// dropped, so no load or store sits in a branch (a bounds branch around each load forces a wait for every load in
// flight at the join, which would serialise the look-ahead). A lane's (x, y) position gets the byte offset of its row in
// a plane, or BUF_OOB when (x, y) lies outside the box; plane z adds z * plane_bytes in 32-bit unsigned arithmetic, so
// z < 0 wraps to >= 2^31 and z >= n lands at or past the end. Both stay out of range while
// (N + 4 n^2) * sizeof(T) < 2^31, which the host launchers check.
constexpr unsigned BUF_OOB = 0x80000000u;
typedef __amdgpu_buffer_rsrc_t BufRes;
__device__ __forceinline__ BufRes field_res(const void *base, long bytes) {
    // word 3 = 0x00020000: 32-bit data format, no swizzle, raw addressing (the gfx9 family's value)
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ unsigned row_off(int xx, int yy, int n, int esize) {
    return (xx >= 0 && xx < n && yy >= 0 && yy < n) ? (unsigned)(xx + n * yy) * (unsigned)esize : BUF_OOB;
}
template <typename T> __device__ __forceinline__ T bload(BufRes r, unsigned off);
template <> __device__ __forceinline__ float bload<float>(BufRes r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double bload<double>(BufRes r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
// aux 2 = nontemporal: the stencil output streams out and should not evict the inputs still in use
template <typename T> __device__ __forceinline__ void bstore_nt(T v, BufRes r, unsigned off);
template <> __device__ __forceinline__ void bstore_nt<float>(float v, BufRes r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 2);
}
template <> __device__ __forceinline__ void bstore_nt<double>(double v, BufRes r, unsigned off) {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, off, 0, 2);
}

// ------------------------------------------------------------------------------------------ stencil ----
// 2.5-D z-streaming: a workgroup owns a TX x TY tile of (x, y) columns of one band and a chunk of z planes, TX = 64
// (one wave per row) or 32 (a wave covers two rows: for n just above a multiple of 32, e.g. 95 = 3 x 32 - 1, the 64-wide
// tile would leave a quarter of its lanes outside the box). Each thread keeps its RPT columns' 9-plane z-window in
// registers; the x/y taps of the current plane come from an LDS tile with a 4-point halo (double-buffered: one barrier
// per plane). The potential V is read once per point.

// Register state of one thread: the 9-plane z window of its RPT columns (plane p in slot (p - z0 + 4) mod 9) and the
// look-ahead queues of its halo cells, V and R (plane p in slot (p - z0) mod 3). The z loop is unrolled by 9 so every
// slot index is a compile-time constant: registers rotate by name and are never copied, so a look-ahead load is only
// waited for when its plane is used.
// XCD-aware workgroup order. The dispatcher hands consecutive workgroup ids to the 8 XCDs round-robin (id mod 8);
// each XCD has its own L2. The stencil launches a 1-D grid and gives XCD x the contiguous logical range
// [x P, (x + 1) P) of (band, z chunk, tile) work items, tiles fastest: the tiles of one band (whose halos overlap) and
// neighbouring bands (which share V) then run together on one XCD and meet in its L2.
struct StWork {
    int tile_x, tile_y, band, zchunk;
    bool valid;
};
// order 0: tiles fastest, then z chunks, then bands; 1: bands fastest; G >= 2: bands in groups of G fastest, then
// tiles, then z chunks, then groups — the G bands of a tile share V in L2 while neighbouring tiles (the halos) run
// side by side. Bands padded to a multiple of G are skipped.
__device__ __forceinline__ StWork st_decode(int ntx, int nty, int nzc, int nb, int order) {
    const int tiles = ntx * nty;
    const int nbp = order >= 2 ? ((nb + order - 1) / order) * order : nb;
    const int total = tiles * nzc * nbp;
    const int per = (total + 7) / 8;
    const int lin = blockIdx.x;
    const int L = (lin & 7) * per + (lin >> 3);
    StWork w;
    w.valid = L < total;
    int tile, rest;
    if (order == 0) {
        tile = L % tiles;
        rest = L / tiles;
        w.zchunk = rest % nzc;
        w.band = rest / nzc;
    } else if (order == 1) {
        w.band = L % nb;
        rest = L / nb;
        tile = rest % tiles;
        w.zchunk = rest / tiles;
    } else {
        const int gi = L % order;
        rest = L / order;
        tile = rest % tiles;
        rest = rest / tiles;
        w.zchunk = rest % nzc;
        w.band = (rest / nzc) * order + gi;
        w.valid = w.valid && w.band < nb;
    }
    w.tile_x = tile % ntx;
    w.tile_y = tile / ntx;
    return w;
}
static inline unsigned st_grid(int ntx, int nty, int nzc, int nb, int order) {
    const long nbp = order >= 2 ? ((nb + order - 1) / order) * order : nb;
    const long total = (long)ntx * nty * nzc * nbp;
    return (unsigned)(8 * ((total + 7) / 8));
}

// Compile-time plane unrolling: f(integral_constant<int, S>) for S = 0 .. W-1 (every register slot index a constant)
template <int S, int W> struct PlaneUnroll {
    template <typename F> __device__ __forceinline__ static void run(F &f) {
        f(std::integral_constant<int, S>{});
        PlaneUnroll<S + 1, W>::run(f);
    }
};
template <int W> struct PlaneUnroll<W, W> {
    template <typename F> __device__ __forceinline__ static void run(F &) {}
};

// The z window holds W >= 9 planes: the 9 the 8th-order z taps of the current plane z read (z - 4 .. z + 4) and
// W - 9 planes of look-ahead, so the plane a step loads (z + W - 5, into the slot plane z - 5 frees) is first used
// W - 9 steps later. W is a multiple of 3, the length of the halo / V / R queues (loaded two planes ahead).
template <typename T, int RPT, int W> struct StRegs {
    T q[W][RPT];
    T hq[3][3];
    T vq[3][RPT];
    T rq[3][RPT];
};

template <typename T, int RPT> struct StCtx {
    BufRes bx, bo, bv, br;  // x and out of this band, V, R of this band
    unsigned pb;            // bytes per plane
    unsigned co[RPT];       // this lane's centre rows (gx, y0 + ty + (256 / TX) rr): byte offsets in a plane, or BUF_OOB
    unsigned ho0, ho1, hox; // its y-halo cells (gx, y0 - 4 + hly0 / hly1; hly1 only at TX = 64) and x-halo cell
    int n;
    int gx, x0, y0, tx, ty;
    int hly0, hly1, hxr, hlx;
    bool has_xh;
    T inv_h2, scale, omega, diag_t, lamb;
    __device__ T ld(BufRes r, unsigned off, int zz) const { return bload<T>(r, off + (unsigned)zz * pb); }
};

template <typename T, int MODE, int RPT, int W, int TX, int SLOT>
__device__ __forceinline__ void st_side(StRegs<T, RPT, W> &R, const StCtx<T, RPT> &c, int zz) {
    constexpr int QS = SLOT % 3;
    R.hq[QS][0] = c.ld(c.bx, c.ho0, zz);
    if (TX == 64) R.hq[QS][1] = c.ld(c.bx, c.ho1, zz);
    R.hq[QS][2] = c.ld(c.bx, c.hox, zz);
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) {
        if (MODE != ST_LAP) R.vq[QS][rr] = c.ld(c.bv, c.co[rr], zz);
        if (MODE == ST_JAC) R.rq[QS][rr] = c.ld(c.br, c.co[rr], zz);
    }
}

// one plane z of the stream; S = (z - z0) mod W
template <typename T, int MODE, int RPT, int W, int TX, int S>
__device__ __forceinline__ void st_plane(StRegs<T, RPT, W> &R, const StCtx<T, RPT> &c,
                                         T (*tile)[(256 / TX) * RPT + 8][TX + 8], int z, int buf) {
    constexpr int TYB = 256 / TX;  // rows per pass of the workgroup
    constexpr int C0 = (S + 4) % W;
    constexpr int NEW = (S + W - 1) % W;
    constexpr int HS = S % 3;
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) R.q[NEW][rr] = c.ld(c.bx, c.co[rr], z + W - 5);
    st_side<T, MODE, RPT, W, TX, S + 2>(R, c, z + 2);
    T(*tl)[TX + 8] = tile[buf];
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) tl[c.ty + TYB * rr + 4][c.tx + 4] = R.q[C0][rr];
    tl[c.hly0][c.tx + 4] = R.hq[HS][0];
    if (TX == 64) tl[c.hly1][c.tx + 4] = R.hq[HS][1];
    if (c.has_xh) tl[c.hxr][c.hlx] = R.hq[HS][2];
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) {
        const int ly = c.ty + TYB * rr + 4, lx = c.tx + 4;
        const T ce = R.q[C0][rr];
        T ax = T(C1) * ((tl[ly][lx + 1] - ce) + (tl[ly][lx - 1] - ce)) + T(C2) * ((tl[ly][lx + 2] - ce) + (tl[ly][lx - 2] - ce)) +
               T(C3) * ((tl[ly][lx + 3] - ce) + (tl[ly][lx - 3] - ce)) + T(C4) * ((tl[ly][lx + 4] - ce) + (tl[ly][lx - 4] - ce));
        T ay = T(C1) * ((tl[ly + 1][lx] - ce) + (tl[ly - 1][lx] - ce)) + T(C2) * ((tl[ly + 2][lx] - ce) + (tl[ly - 2][lx] - ce)) +
               T(C3) * ((tl[ly + 3][lx] - ce) + (tl[ly - 3][lx] - ce)) + T(C4) * ((tl[ly + 4][lx] - ce) + (tl[ly - 4][lx] - ce));
        T az = T(C1) * ((R.q[(S + 5) % W][rr] - ce) + (R.q[(S + 3) % W][rr] - ce)) +
               T(C2) * ((R.q[(S + 6) % W][rr] - ce) + (R.q[(S + 2) % W][rr] - ce)) +
               T(C3) * ((R.q[(S + 7) % W][rr] - ce) + (R.q[(S + 1) % W][rr] - ce)) +
               T(C4) * ((R.q[(S + 8) % W][rr] - ce) + (R.q[S % W][rr] - ce));
        const T lap = ((ax + ay) + az) * c.inv_h2;
        T o;
        if (MODE == ST_LAP) {
            o = c.scale * lap;
        } else if (MODE == ST_HAM) {
            o = T(-0.5) * lap + R.vq[HS][rr] * ce;
        } else {
            const T vv = R.vq[HS][rr];
            const T hw = T(-0.5) * lap + vv * ce - c.lamb * ce;
            T d = c.diag_t + vv - c.lamb;
            d = d > T(0.1) ? d : T(0.1);
            o = ce + c.omega / d * (R.rq[HS][rr] - hw);
        }
        // a row outside the box has offset BUF_OOB: its store is dropped
        bstore_nt<T>(o, c.bo, c.co[rr] + (unsigned)z * c.pb);
    }
}

template <typename T, int MODE, int RPT, int W, int TX>
__global__ __launch_bounds__(256) void k_stencil(const T *__restrict__ x, T *__restrict__ out, const T *__restrict__ v,
                                                 const T *__restrict__ r, const double *__restrict__ lam, int n,
                                                 long N, int zc, int nb, int order, T inv_h2, T scale, T omega,
                                                 T diag_t, const T *__restrict__ x2, T *__restrict__ out2, int nb1) {
    static_assert(TX == 64 || TX == 32, "tile width");
    constexpr int TYB = 256 / TX;  // workgroup (TX, TYB)
    constexpr int TY = TYB * RPT;
    // the y halo (8 rows of TX cells) takes 512 / TX cells per lane, the x halo (TY + 8 rows of 8 cells) one
    static_assert(8 * (TY + 8) <= 256, "x halo larger than the workgroup");
    __shared__ T tile[2][TY + 8][TX + 8];
    StCtx<T, RPT> c;
    c.tx = threadIdx.x;
    c.ty = threadIdx.y;
    const int t = c.ty * TX + c.tx;
    const int nzc = (n + zc - 1) / zc;
    const StWork wk = st_decode((n + TX - 1) / TX, (n + TY - 1) / TY, nzc, nb, order);
    if (!wk.valid) return;
    c.x0 = wk.tile_x * TX;
    c.y0 = wk.tile_y * TY;
    const int band = wk.band;
    const int z0 = wk.zchunk * zc;
    const int z1 = min(n, z0 + zc);
    const long fb = N * (long)sizeof(T);
    c.bx = field_res(band < nb1 ? x + (long)band * N : x2 + (long)(band - nb1) * N, fb);
    c.bo = field_res(band < nb1 ? out + (long)band * N : out2 + (long)(band - nb1) * N, fb);
    c.bv = field_res(v, MODE == ST_LAP ? 0 : fb);
    c.br = field_res(MODE == ST_JAC ? r + (long)band * N : x, MODE == ST_JAC ? fb : 0);
    c.lamb = MODE == ST_JAC ? (T)lam[band] : T(0);
    c.n = n;
    c.pb = (unsigned)((long)n * n * (long)sizeof(T));
    c.gx = c.x0 + c.tx;
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) c.co[rr] = row_off(c.gx, c.y0 + c.ty + TYB * rr, n, (int)sizeof(T));
    // y-halo rows 0..3 above the tile, 4..7 below it (LDS rows TY + 4 ..): row ty, and at TX = 64 also row ty + 4
    const int hy0 = c.ty, hy1 = c.ty + 4;
    c.hly0 = hy0 < 4 ? hy0 : TY + hy0;
    c.hly1 = TY + hy1;
    c.has_xh = t < 8 * (TY + 8);
    c.hxr = t >> 3;
    const int hxc = t & 7;
    c.hlx = hxc < 4 ? hxc : TX + hxc;
    c.ho0 = row_off(c.gx, c.y0 - 4 + c.hly0, n, (int)sizeof(T));
    c.ho1 = TX == 64 ? row_off(c.gx, c.y0 - 4 + c.hly1, n, (int)sizeof(T)) : BUF_OOB;
    c.hox = c.has_xh ? row_off(c.x0 - 4 + c.hlx, c.y0 - 4 + c.hxr, n, (int)sizeof(T)) : BUF_OOB;
    c.inv_h2 = inv_h2;
    c.scale = scale;
    c.omega = omega;
    c.diag_t = diag_t;

    StRegs<T, RPT, W> R;
#pragma unroll
    for (int d = 0; d < W - 1; d++)
#pragma unroll
        for (int rr = 0; rr < RPT; rr++) R.q[d][rr] = c.ld(c.bx, c.co[rr], z0 - 4 + d);
    st_side<T, MODE, RPT, W, TX, 0>(R, c, z0);
    st_side<T, MODE, RPT, W, TX, 1>(R, c, z0 + 1);

    // consecutive planes alternate the two LDS buffers (plane p uses buffer (p - z0) & 1): one barrier per plane then
    // separates every write of a buffer from the previous reads of it
    int z = z0;
    for (; z + W <= z1; z += W) {
        const int b = (z - z0) & 1;
        auto step = [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            st_plane<T, MODE, RPT, W, TX, S>(R, c, tile, z + S, b ^ (S & 1));
        };
        PlaneUnroll<0, W>::run(step);
    }
    // remainder (< W planes) with the same slot sequence
    const int b = (z - z0) & 1;
    auto tail = [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        if (z + S < z1) st_plane<T, MODE, RPT, W, TX, S>(R, c, tile, z + S, b ^ (S & 1));
    };
    PlaneUnroll<0, W - 1>::run(tail);
}

// ---- vectorised stencil: each lane owns VX consecutive x points (16-byte loads and stores: float4 / double2) ----
// Requires n % VX == 0 so every vector is aligned and lies wholly inside or outside the box. Workgroup 16 x 16 lanes:
// tile TX = 16 VX columns x TY = 16 RPT rows (row ty + 16 rr). z window, halo/V/R look-ahead and LDS double buffering
// as the scalar kernel; x taps are read from LDS as three 16-byte vectors, y taps as one vector per row.
template <typename T> struct Vec16;
template <> struct Vec16<float> {
    typedef float4 type;
    static constexpr int N = 4;
};
template <> struct Vec16<double> {
    typedef double2 type;
    static constexpr int N = 2;
};

template <typename T, int RPT, int W> struct SvRegs {
    typedef typename Vec16<T>::type V;
    V q[W][RPT];
    V hq[3];
    V hq2[3];  // second x-halo vector when a vector holds fewer than the 4 halo points (double2)
    V vq[3][RPT];
    V rq[3][RPT];
};

// native 16-byte vectors for the buffer intrinsics
typedef float lira_nf4 __attribute__((ext_vector_type(4)));
typedef double lira_nd2 __attribute__((ext_vector_type(2)));

template <typename T> __device__ __forceinline__ T vget(const typename Vec16<T>::type &v, int i);
template <> __device__ __forceinline__ float vget<float>(const float4 &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}
template <> __device__ __forceinline__ double vget<double>(const double2 &v, int i) { return i == 0 ? v.x : v.y; }
template <typename T> __device__ __forceinline__ void vset(typename Vec16<T>::type &v, int i, T x);
template <> __device__ __forceinline__ void vset<float>(float4 &v, int i, float x) {
    if (i == 0) v.x = x; else if (i == 1) v.y = x; else if (i == 2) v.z = x; else v.w = x;
}
template <> __device__ __forceinline__ void vset<double>(double2 &v, int i, double x) {
    if (i == 0) v.x = x; else v.y = x;
}
// 16-byte buffer load / nontemporal store of one vector (float4 or double2)
template <typename T> __device__ __forceinline__ typename Vec16<T>::type bloadv(BufRes r, unsigned off);
template <> __device__ __forceinline__ float4 bloadv<float>(BufRes r, unsigned off) {
    const lira_nf4 t = __builtin_bit_cast(lira_nf4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    return make_float4(t.x, t.y, t.z, t.w);
}
template <> __device__ __forceinline__ double2 bloadv<double>(BufRes r, unsigned off) {
    const lira_nd2 t = __builtin_bit_cast(lira_nd2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    return make_double2(t.x, t.y);
}
__device__ __forceinline__ void bstorev_nt(const float4 &v, BufRes r, unsigned off) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const lira_nf4 t = {v.x, v.y, v.z, v.w};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, t), r, off, 0, 2);
}
__device__ __forceinline__ void bstorev_nt(const double2 &v, BufRes r, unsigned off) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const lira_nd2 t = {v.x, v.y};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, t), r, off, 0, 2);
}

template <typename T, int RPT> struct SvCtx {
    typedef typename Vec16<T>::type V;
    static constexpr int VX = Vec16<T>::N;
    BufRes bx, bo, bv, br;
    unsigned pb;                // bytes per plane
    unsigned co[RPT];           // centre rows (gx, y0 + ty + 16 rr): byte offsets in a plane, or BUF_OOB
    unsigned ho, ho2;           // this lane's halo vector(s), or BUF_OOB
    int n;
    int x0, y0, tx, ty, gx;
    int h_ly, h_lx;             // this lane's halo vector (LDS row/col), if any
    bool has_h, has_h2;         // has_h2: an x-halo lane that also loads the vector at h_gx + VX
    T inv_h2, scale, omega, diag_t, lamb;
    __device__ V ldv(BufRes r, unsigned off, int zz) const { return bloadv<T>(r, off + (unsigned)zz * pb); }
};

template <typename T, int MODE, int RPT, int W, int SLOT>
__device__ __forceinline__ void sv_side(SvRegs<T, RPT, W> &R, const SvCtx<T, RPT> &c, int zz) {
    constexpr int QS = SLOT % 3;
    R.hq[QS] = c.ldv(c.bx, c.ho, zz);
    if (SvCtx<T, RPT>::VX < 4) R.hq2[QS] = c.ldv(c.bx, c.ho2, zz);
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) {
        if (MODE != ST_LAP) R.vq[QS][rr] = c.ldv(c.bv, c.co[rr], zz);
        if (MODE == ST_JAC) R.rq[QS][rr] = c.ldv(c.br, c.co[rr], zz);
    }
}

template <typename T, int MODE, int RPT, int W, int S, int LXP>
__device__ __forceinline__ void sv_plane(SvRegs<T, RPT, W> &R, const SvCtx<T, RPT> &c, T (*tile)[16 * RPT + 8][LXP],
                                         int z, int buf) {
    typedef typename Vec16<T>::type V;
    constexpr int VX = Vec16<T>::N;
    constexpr int C0 = (S + 4) % W;
    constexpr int NEW = (S + W - 1) % W;
    constexpr int HS = S % 3;
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) R.q[NEW][rr] = c.ldv(c.bx, c.co[rr], z + W - 5);
    sv_side<T, MODE, RPT, W, S + 2>(R, c, z + 2);
    T(*tl)[LXP] = tile[buf];
    const int lx = 4 + VX * c.tx;  // LDS column of this lane's first point
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) *reinterpret_cast<V *>(&tl[c.ty + 16 * rr + 4][lx]) = R.q[C0][rr];
    if (c.has_h) *reinterpret_cast<V *>(&tl[c.h_ly][c.h_lx]) = R.hq[HS];
    if (VX < 4 && c.has_h2) *reinterpret_cast<V *>(&tl[c.h_ly][c.h_lx + VX]) = R.hq2[HS];
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) {
        const int ly = c.ty + 16 * rr + 4;
        // x window: 4 + VX + 4 values starting at lx - 4
        T w[VX + 8];
#pragma unroll
        for (int b = 0; b < (VX + 8) / VX; b++) {
            const V vv = *reinterpret_cast<const V *>(&tl[ly][lx - 4 + VX * b]);
#pragma unroll
            for (int i = 0; i < VX; i++) w[VX * b + i] = vget<T>(vv, i);
        }
        V ym[4], yp[4];
#pragma unroll
        for (int m = 1; m <= 4; m++) {
            ym[m - 1] = *reinterpret_cast<const V *>(&tl[ly - m][lx]);
            yp[m - 1] = *reinterpret_cast<const V *>(&tl[ly + m][lx]);
        }
        V outv;
#pragma unroll
        for (int i = 0; i < VX; i++) {
            const T ce = w[4 + i];
            const T ax = T(C1) * ((w[5 + i] - ce) + (w[3 + i] - ce)) + T(C2) * ((w[6 + i] - ce) + (w[2 + i] - ce)) +
                         T(C3) * ((w[7 + i] - ce) + (w[1 + i] - ce)) + T(C4) * ((w[8 + i] - ce) + (w[0 + i] - ce));
            const T ay = T(C1) * ((vget<T>(yp[0], i) - ce) + (vget<T>(ym[0], i) - ce)) +
                         T(C2) * ((vget<T>(yp[1], i) - ce) + (vget<T>(ym[1], i) - ce)) +
                         T(C3) * ((vget<T>(yp[2], i) - ce) + (vget<T>(ym[2], i) - ce)) +
                         T(C4) * ((vget<T>(yp[3], i) - ce) + (vget<T>(ym[3], i) - ce));
            const T az = T(C1) * ((vget<T>(R.q[(S + 5) % W][rr], i) - ce) + (vget<T>(R.q[(S + 3) % W][rr], i) - ce)) +
                         T(C2) * ((vget<T>(R.q[(S + 6) % W][rr], i) - ce) + (vget<T>(R.q[(S + 2) % W][rr], i) - ce)) +
                         T(C3) * ((vget<T>(R.q[(S + 7) % W][rr], i) - ce) + (vget<T>(R.q[(S + 1) % W][rr], i) - ce)) +
                         T(C4) * ((vget<T>(R.q[(S + 8) % W][rr], i) - ce) + (vget<T>(R.q[S % W][rr], i) - ce));
            const T lap = ((ax + ay) + az) * c.inv_h2;
            T o;
            if (MODE == ST_LAP) {
                o = c.scale * lap;
            } else if (MODE == ST_HAM) {
                o = T(-0.5) * lap + vget<T>(R.vq[HS][rr], i) * ce;
            } else {
                const T vv = vget<T>(R.vq[HS][rr], i);
                const T hw = T(-0.5) * lap + vv * ce - c.lamb * ce;
                T d = c.diag_t + vv - c.lamb;
                d = d > T(0.1) ? d : T(0.1);
                o = ce + c.omega / d * (vget<T>(R.rq[HS][rr], i) - hw);
            }
            vset<T>(outv, i, o);
        }
        // a row or vector outside the box has offset BUF_OOB: its store is dropped
        bstorev_nt(outv, c.bo, c.co[rr] + (unsigned)z * c.pb);
    }
}

template <typename T, int MODE, int RPT, int W>
__global__ __launch_bounds__(256) void k_stencil_vec(const T *__restrict__ x, T *__restrict__ out,
                                                     const T *__restrict__ v, const T *__restrict__ r,
                                                     const double *__restrict__ lam, int n, long N, int zc, int nb,
                                                     int order, T inv_h2, T scale, T omega, T diag_t,
                                                     const T *__restrict__ x2, T *__restrict__ out2, int nb1) {
    constexpr int VX = Vec16<T>::N;
    constexpr int TX = 16 * VX, TY = 16 * RPT;
    constexpr int LXP = TX + 8;
    __shared__ __align__(16) T tile[2][TY + 8][LXP];
    SvCtx<T, RPT> c;
    c.tx = threadIdx.x;
    c.ty = threadIdx.y;
    const int t = c.ty * 16 + c.tx;
    const int nzc = (n + zc - 1) / zc;
    const StWork wk = st_decode((n + TX - 1) / TX, (n + TY - 1) / TY, nzc, nb, order);
    if (!wk.valid) return;
    c.x0 = wk.tile_x * TX;
    c.y0 = wk.tile_y * TY;
    const int band = wk.band;
    const int z0 = wk.zchunk * zc;
    const int z1 = min(n, z0 + zc);
    const long fb = N * (long)sizeof(T);
    c.bx = field_res(band < nb1 ? x + (long)band * N : x2 + (long)(band - nb1) * N, fb);
    c.bo = field_res(band < nb1 ? out + (long)band * N : out2 + (long)(band - nb1) * N, fb);
    c.bv = field_res(v, MODE == ST_LAP ? 0 : fb);
    c.br = field_res(MODE == ST_JAC ? r + (long)band * N : x, MODE == ST_JAC ? fb : 0);
    c.lamb = MODE == ST_JAC ? (T)lam[band] : T(0);
    c.n = n;
    c.pb = (unsigned)((long)n * n * (long)sizeof(T));
    c.gx = c.x0 + VX * c.tx;
#pragma unroll
    for (int rr = 0; rr < RPT; rr++) c.co[rr] = row_off(c.gx, c.y0 + c.ty + 16 * rr, n, (int)sizeof(T));
    int h_gx = 0, h_gy = 0;
    // halo vectors: 8 y-halo rows x 16 vectors (lanes 0..127), then (TY + 8) rows x 2 sides (lanes 128..); a side
    // holds 4 points, i.e. 4 / VX vectors per x-halo lane
    if (t < 128) {
        const int row = t >> 4, col = t & 15;
        c.h_ly = row < 4 ? row : TY + row;
        c.h_lx = 4 + VX * col;
        h_gx = c.x0 + VX * col;
        h_gy = c.y0 - 4 + c.h_ly;
        c.has_h = true;
    } else if (t < 128 + 2 * (TY + 8)) {
        const int e = t - 128, row = e >> 1, side = e & 1;
        c.h_ly = row;
        c.h_lx = side ? 4 + TX : 0;
        h_gx = side ? c.x0 + TX : c.x0 - 4;
        h_gy = c.y0 - 4 + row;
        c.has_h = true;
    } else {
        c.h_ly = c.h_lx = 0;
        c.has_h = false;
    }
    c.has_h2 = c.has_h && t >= 128;
    c.ho = c.has_h ? row_off(h_gx, h_gy, n, (int)sizeof(T)) : BUF_OOB;
    c.ho2 = c.has_h2 ? row_off(h_gx + VX, h_gy, n, (int)sizeof(T)) : BUF_OOB;
    c.inv_h2 = inv_h2;
    c.scale = scale;
    c.omega = omega;
    c.diag_t = diag_t;
    SvRegs<T, RPT, W> R;
#pragma unroll
    for (int d = 0; d < W - 1; d++)
#pragma unroll
        for (int rr = 0; rr < RPT; rr++) R.q[d][rr] = c.ldv(c.bx, c.co[rr], z0 - 4 + d);
    sv_side<T, MODE, RPT, W, 0>(R, c, z0);
    sv_side<T, MODE, RPT, W, 1>(R, c, z0 + 1);
    int z = z0;
    for (; z + W <= z1; z += W) {
        const int b = (z - z0) & 1;
        auto step = [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            sv_plane<T, MODE, RPT, W, S, LXP>(R, c, tile, z + S, b ^ (S & 1));
        };
        PlaneUnroll<0, W>::run(step);
    }
    const int b = (z - z0) & 1;
    auto tail = [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        if (z + S < z1) sv_plane<T, MODE, RPT, W, S, LXP>(R, c, tile, z + S, b ^ (S & 1));
    };
    PlaneUnroll<0, W - 1>::run(tail);
}

// ---- vector lanes carrying NB bands: lane (tx, ty) owns VX x points of row y0 + ty for bands b0 .. b0 + NB - 1 ----
// V (and R in ST_JAC per band) is loaded once per point for the NB bands; each band has its own LDS plane.
template <typename T, int NB, int W> struct VbRegs {
    typedef typename Vec16<T>::type V;
    V q[W][NB];
    V hq[3][NB];
    V hq2[3][NB];
    V vq[3];
    V rq[3][NB];
};

template <typename T, int NB> struct VbCtx {
    typedef typename Vec16<T>::type V;
    static constexpr int VX = Vec16<T>::N;
    BufRes bx[NB], bo[NB], br[NB], bv;
    T lamb[NB];
    unsigned so[NB];   // this lane's store offset per band: co, or BUF_OOB for a band past the block
    unsigned pb;       // bytes per plane
    unsigned co;       // centre row (gx, gy): byte offset in a plane, or BUF_OOB
    unsigned ho, ho2;  // halo vector(s), or BUF_OOB
    int n;
    int x0, y0, tx, ty, gx, gy;
    int h_ly, h_lx;
    bool has_h, has_h2;
    T inv_h2, scale, omega, diag_t;
    __device__ V ldv(BufRes r, unsigned off, int zz) const { return bloadv<T>(r, off + (unsigned)zz * pb); }
};

template <typename T, int MODE, int NB, int W, int SLOT>
__device__ __forceinline__ void vb_side(VbRegs<T, NB, W> &R, const VbCtx<T, NB> &c, int zz) {
    constexpr int QS = SLOT % 3;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        R.hq[QS][b] = c.ldv(c.bx[b], c.ho, zz);
        if (VbCtx<T, NB>::VX < 4) R.hq2[QS][b] = c.ldv(c.bx[b], c.ho2, zz);
        if (MODE == ST_JAC) R.rq[QS][b] = c.ldv(c.br[b], c.co, zz);
    }
    if (MODE != ST_LAP) R.vq[QS] = c.ldv(c.bv, c.co, zz);
}

template <typename T, int MODE, int NB, int W, int S, int TYP, int LXP>
__device__ __forceinline__ void vb_plane(VbRegs<T, NB, W> &R, const VbCtx<T, NB> &c, T (*tile)[NB][TYP][LXP], int z,
                                         int buf) {
    typedef typename Vec16<T>::type V;
    constexpr int VX = Vec16<T>::N;
    constexpr int C0 = (S + 4) % W;
    constexpr int NEW = (S + W - 1) % W;
    constexpr int HS = S % 3;
#pragma unroll
    for (int b = 0; b < NB; b++) R.q[NEW][b] = c.ldv(c.bx[b], c.co, z + W - 5);
    vb_side<T, MODE, NB, W, S + 2>(R, c, z + 2);
    const int lx = 4 + VX * c.tx, ly = c.ty + 4;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        *reinterpret_cast<V *>(&tile[buf][b][ly][lx]) = R.q[C0][b];
        if (c.has_h) *reinterpret_cast<V *>(&tile[buf][b][c.h_ly][c.h_lx]) = R.hq[HS][b];
        if (VX < 4 && c.has_h2) *reinterpret_cast<V *>(&tile[buf][b][c.h_ly][c.h_lx + VX]) = R.hq2[HS][b];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; b++) {
        T(*tl)[LXP] = tile[buf][b];
        T w[VX + 8];
#pragma unroll
        for (int q = 0; q < (VX + 8) / VX; q++) {
            const V vv = *reinterpret_cast<const V *>(&tl[ly][lx - 4 + VX * q]);
#pragma unroll
            for (int i = 0; i < VX; i++) w[VX * q + i] = vget<T>(vv, i);
        }
        V ym[4], yp[4];
#pragma unroll
        for (int m = 1; m <= 4; m++) {
            ym[m - 1] = *reinterpret_cast<const V *>(&tl[ly - m][lx]);
            yp[m - 1] = *reinterpret_cast<const V *>(&tl[ly + m][lx]);
        }
        V outv;
#pragma unroll
        for (int i = 0; i < VX; i++) {
            const T ce = w[4 + i];
            const T ax = T(C1) * ((w[5 + i] - ce) + (w[3 + i] - ce)) + T(C2) * ((w[6 + i] - ce) + (w[2 + i] - ce)) +
                         T(C3) * ((w[7 + i] - ce) + (w[1 + i] - ce)) + T(C4) * ((w[8 + i] - ce) + (w[0 + i] - ce));
            const T ay = T(C1) * ((vget<T>(yp[0], i) - ce) + (vget<T>(ym[0], i) - ce)) +
                         T(C2) * ((vget<T>(yp[1], i) - ce) + (vget<T>(ym[1], i) - ce)) +
                         T(C3) * ((vget<T>(yp[2], i) - ce) + (vget<T>(ym[2], i) - ce)) +
                         T(C4) * ((vget<T>(yp[3], i) - ce) + (vget<T>(ym[3], i) - ce));
            const T az = T(C1) * ((vget<T>(R.q[(S + 5) % W][b], i) - ce) + (vget<T>(R.q[(S + 3) % W][b], i) - ce)) +
                         T(C2) * ((vget<T>(R.q[(S + 6) % W][b], i) - ce) + (vget<T>(R.q[(S + 2) % W][b], i) - ce)) +
                         T(C3) * ((vget<T>(R.q[(S + 7) % W][b], i) - ce) + (vget<T>(R.q[(S + 1) % W][b], i) - ce)) +
                         T(C4) * ((vget<T>(R.q[(S + 8) % W][b], i) - ce) + (vget<T>(R.q[S % W][b], i) - ce));
            const T lap = ((ax + ay) + az) * c.inv_h2;
            T o;
            if (MODE == ST_LAP) {
                o = c.scale * lap;
            } else if (MODE == ST_HAM) {
                o = T(-0.5) * lap + vget<T>(R.vq[HS], i) * ce;
            } else {
                const T vv = vget<T>(R.vq[HS], i);
                const T hw = T(-0.5) * lap + vv * ce - c.lamb[b] * ce;
                T d = c.diag_t + vv - c.lamb[b];
                d = d > T(0.1) ? d : T(0.1);
                o = ce + c.omega / d * (vget<T>(R.rq[HS][b], i) - hw);
            }
            vset<T>(outv, i, o);
        }
        bstorev_nt(outv, c.bo[b], c.so[b] + (unsigned)z * c.pb);
    }
}

template <typename T, int MODE, int NB, int W>
__global__ __launch_bounds__(256) void k_stencil_vb(const T *__restrict__ x, T *__restrict__ out,
                                                    const T *__restrict__ v, const T *__restrict__ r,
                                                    const double *__restrict__ lam, int n, long N, int zc, int nb,
                                                    int order, T inv_h2, T scale, T omega, T diag_t,
                                                    const T *__restrict__ x2, T *__restrict__ out2, int nb1) {
    constexpr int VX = Vec16<T>::N;
    constexpr int TX = 16 * VX, TY = 16;
    constexpr int LXP = TX + 8, TYP = TY + 8;
    __shared__ __align__(16) T tile[2][NB][TYP][LXP];
    VbCtx<T, NB> c;
    c.tx = threadIdx.x;
    c.ty = threadIdx.y;
    const int t = c.ty * 16 + c.tx;
    const int nzc = (n + zc - 1) / zc;
    const int ngroups = (nb + NB - 1) / NB;
    const StWork wk = st_decode((n + TX - 1) / TX, (n + TY - 1) / TY, nzc, ngroups, order);
    if (!wk.valid) return;
    c.x0 = wk.tile_x * TX;
    c.y0 = wk.tile_y * TY;
    const int z0 = wk.zchunk * zc;
    const int z1 = min(n, z0 + zc);
    const long fb = N * (long)sizeof(T);
    c.n = n;
    c.pb = (unsigned)((long)n * n * (long)sizeof(T));
    c.gx = c.x0 + VX * c.tx;
    c.gy = c.y0 + c.ty;
    c.co = row_off(c.gx, c.gy, n, (int)sizeof(T));
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int band = wk.band * NB + b;
        const bool bok = band < nb;
        const int bb = bok ? band : wk.band * NB;  // a missing band reads a valid one and stores nothing
        c.bx[b] = field_res(bb < nb1 ? x + (long)bb * N : x2 + (long)(bb - nb1) * N, fb);
        c.bo[b] = field_res(bb < nb1 ? out + (long)bb * N : out2 + (long)(bb - nb1) * N, fb);
        c.br[b] = field_res(MODE == ST_JAC ? r + (long)bb * N : x, MODE == ST_JAC ? fb : 0);
        c.lamb[b] = MODE == ST_JAC ? (T)lam[bb] : T(0);
        c.so[b] = bok ? c.co : BUF_OOB;
    }
    c.bv = field_res(v, MODE == ST_LAP ? 0 : fb);
    int h_gx = 0, h_gy = 0;
    if (t < 128) {
        const int row = t >> 4, col = t & 15;
        c.h_ly = row < 4 ? row : TY + row;
        c.h_lx = 4 + VX * col;
        h_gx = c.x0 + VX * col;
        h_gy = c.y0 - 4 + c.h_ly;
        c.has_h = true;
    } else if (t < 128 + 2 * TYP) {
        const int e = t - 128, row = e >> 1, side = e & 1;
        c.h_ly = row;
        c.h_lx = side ? 4 + TX : 0;
        h_gx = side ? c.x0 + TX : c.x0 - 4;
        h_gy = c.y0 - 4 + row;
        c.has_h = true;
    } else {
        c.h_ly = c.h_lx = 0;
        c.has_h = false;
    }
    c.has_h2 = c.has_h && t >= 128;
    c.ho = c.has_h ? row_off(h_gx, h_gy, n, (int)sizeof(T)) : BUF_OOB;
    c.ho2 = c.has_h2 ? row_off(h_gx + VX, h_gy, n, (int)sizeof(T)) : BUF_OOB;
    c.inv_h2 = inv_h2;
    c.scale = scale;
    c.omega = omega;
    c.diag_t = diag_t;
    VbRegs<T, NB, W> R;
#pragma unroll
    for (int d = 0; d < W - 1; d++)
#pragma unroll
        for (int b = 0; b < NB; b++) R.q[d][b] = c.ldv(c.bx[b], c.co, z0 - 4 + d);
    vb_side<T, MODE, NB, W, 0>(R, c, z0);
    vb_side<T, MODE, NB, W, 1>(R, c, z0 + 1);
    int z = z0;
    for (; z + W <= z1; z += W) {
        const int b = (z - z0) & 1;
        auto step = [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            vb_plane<T, MODE, NB, W, S, TYP, LXP>(R, c, tile, z + S, b ^ (S & 1));
        };
        PlaneUnroll<0, W>::run(step);
    }
    const int b = (z - z0) & 1;
    auto tail = [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        if (z + S < z1) vb_plane<T, MODE, NB, W, S, TYP, LXP>(R, c, tile, z + S, b ^ (S & 1));
    };
    PlaneUnroll<0, W - 1>::run(tail);
}

struct StVariant {
    int rpt;
    int zc;     // 0 = whole column
    int vec;    // 1: 16-byte vector lanes (needs n divisible by the vector width); 2: lanes carrying rpt bands;
                // 3: scalar lanes on a 32-wide tile (rpt 1 or 2)
    int order;  // XCD work order: 0 tiles fastest, 1 bands fastest, G >= 2 bands in groups of G fastest
    int win;    // z window: 9 (no look-ahead beyond the taps) or 12 (three planes of look-ahead)
    const char *name;
};
static const StVariant ST_VARIANTS[] = {
    {4, 0, 0, 0, 9, "64x16 tile, full z column"},
    {4, 32, 0, 0, 9, "64x16 tile, z chunks of 32"},
    {4, 16, 0, 0, 9, "64x16 tile, z chunks of 16"},
    {2, 0, 0, 0, 9, "64x8 tile, full z column"},
    {2, 16, 0, 0, 9, "64x8 tile, z chunks of 16"},
    {2, 8, 0, 0, 9, "64x8 tile, z chunks of 8"},
    {1, 0, 1, 0, 9, "vector lanes, 16 rows, full z column"},
    {1, 32, 1, 0, 9, "vector lanes, 16 rows, z chunks of 32"},
    {1, 16, 1, 0, 9, "vector lanes, 16 rows, z chunks of 16"},
    {2, 0, 1, 0, 9, "vector lanes, 32 rows, full z column"},
    {2, 32, 1, 0, 9, "vector lanes, 32 rows, z chunks of 32"},
    {1, 0, 1, 1, 9, "vector lanes, 16 rows, full z column, bands fastest"},
    {1, 32, 1, 1, 9, "vector lanes, 16 rows, z chunks of 32, bands fastest"},
    {2, 0, 1, 1, 9, "vector lanes, 32 rows, full z column, bands fastest"},
    {2, 0, 0, 1, 9, "64x8 tile, full z column, bands fastest"},
    {2, 16, 0, 1, 9, "64x8 tile, z chunks of 16, bands fastest"},
    {2, 0, 2, 0, 9, "vector lanes, 2 bands per lane, full z column"},
    {2, 32, 2, 0, 9, "vector lanes, 2 bands per lane, z chunks of 32"},
    {2, 0, 2, 1, 9, "vector lanes, 2 bands per lane, full z column, bands fastest"},
    {4, 0, 2, 0, 9, "vector lanes, 4 bands per lane, full z column"},
    {1, 0, 1, 8, 9, "vector lanes, 16 rows, full z column, bands in groups of 8"},
    {1, 0, 1, 4, 9, "vector lanes, 16 rows, full z column, bands in groups of 4"},
    {2, 0, 1, 8, 9, "vector lanes, 32 rows, full z column, bands in groups of 8"},
    {1, 32, 1, 8, 9, "vector lanes, 16 rows, z chunks of 32, bands in groups of 8"},
    {4, 0, 0, 8, 9, "64x16 tile, full z column, bands in groups of 8"},
    {2, 16, 0, 8, 9, "64x8 tile, z chunks of 16, bands in groups of 8"},
    {2, 16, 0, 4, 9, "64x8 tile, z chunks of 16, bands in groups of 4"},
    {4, 0, 0, 0, 12, "64x16 tile, full z column, 12-plane window"},
    {4, 32, 0, 0, 12, "64x16 tile, z chunks of 32, 12-plane window"},
    {2, 0, 0, 0, 12, "64x8 tile, full z column, 12-plane window"},
    {2, 16, 0, 0, 12, "64x8 tile, z chunks of 16, 12-plane window"},
    {2, 16, 0, 4, 12, "64x8 tile, z chunks of 16, bands in groups of 4, 12-plane window"},
    {1, 0, 1, 0, 12, "vector lanes, 16 rows, full z column, 12-plane window"},
    {1, 32, 1, 0, 12, "vector lanes, 16 rows, z chunks of 32, 12-plane window"},
    {1, 0, 1, 4, 12, "vector lanes, 16 rows, full z column, bands in groups of 4, 12-plane window"},
    {2, 0, 1, 0, 12, "vector lanes, 32 rows, full z column, 12-plane window"},
    {2, 0, 2, 0, 12, "vector lanes, 2 bands per lane, full z column, 12-plane window"},
    {2, 0, 3, 0, 9, "32x16 tile, full z column"},
    {2, 16, 3, 0, 9, "32x16 tile, z chunks of 16"},
    {1, 0, 3, 0, 9, "32x8 tile, full z column"},
    {1, 16, 3, 0, 9, "32x8 tile, z chunks of 16"},
    {2, 16, 3, 1, 9, "32x16 tile, z chunks of 16, bands fastest"},
    {2, 16, 3, 4, 9, "32x16 tile, z chunks of 16, bands in groups of 4"},
};
static int g_default_variant[2] = {-1, -1};

int stencil_variants() { return (int)(sizeof(ST_VARIANTS) / sizeof(ST_VARIANTS[0])); }
const char *stencil_variant_name(int v) { return (v >= 0 && v < stencil_variants()) ? ST_VARIANTS[v].name : "auto"; }
void stencil_set_default_variant(int dtype_size, int variant) { g_default_variant[dtype_size == 8 ? 1 : 0] = variant; }

// Heuristic when no tuned choice exists: whole columns when the band x tile count alone fills the chip, z chunks
// otherwise (small grids, few bands).
static int pick_variant(int n, int nb) {
    long tiles16 = (long)((n + 63) / 64) * ((n + 15) / 16) * nb;
    if (tiles16 >= 1024) return 0;
    if (tiles16 * ((n + 31) / 32) >= 1024) return 1;
    return 2;
}

template <typename T, int MODE, int RPT, int W, int TX>
static void launch_stencil(const StencilParams &p, int zc, int order, const T *x, T *out, const T *v, const T *r,
                           const double *lam, hipStream_t s) {
    constexpr int TYB = 256 / TX, TY = TYB * RPT;
    const int n = p.n;
    const long N = (long)n * n * n;
    if (zc <= 0 || zc > n) zc = n;
    const int nzc = (n + zc - 1) / zc;
    const unsigned grid = st_grid((n + TX - 1) / TX, (n + TY - 1) / TY, nzc, p.nb, order);
    dim3 block(TX, TYB, 1);
    hipLaunchKernelGGL((k_stencil<T, MODE, RPT, W, TX>), dim3(grid), block, 0, s, x, out, v, r, lam, n, N, zc, p.nb, order,
                       (T)(1.0 / (p.h * p.h)), (T)p.scale, (T)p.omega, (T)p.diag_t, static_cast<const T *>(p.x2),
                       static_cast<T *>(p.out2), p.nb1 < 0 ? p.nb : p.nb1);
    LIRA_HIP(hipGetLastError());
}

template <typename T, int MODE, int RPT, int W>
static void launch_stencil_vec(const StencilParams &p, int zc, int order, const T *x, T *out, const T *v, const T *r,
                               const double *lam, hipStream_t s) {
    constexpr int VX = Vec16<T>::N;
    const int n = p.n;
    const long N = (long)n * n * n;
    if (zc <= 0 || zc > n) zc = n;
    const int nzc = (n + zc - 1) / zc;
    const unsigned grid = st_grid((n + 16 * VX - 1) / (16 * VX), (n + 16 * RPT - 1) / (16 * RPT), nzc, p.nb, order);
    hipLaunchKernelGGL((k_stencil_vec<T, MODE, RPT, W>), dim3(grid), dim3(16, 16, 1), 0, s, x, out, v, r, lam, n, N, zc,
                       p.nb, order, (T)(1.0 / (p.h * p.h)), (T)p.scale, (T)p.omega, (T)p.diag_t,
                       static_cast<const T *>(p.x2), static_cast<T *>(p.out2), p.nb1 < 0 ? p.nb : p.nb1);
    LIRA_HIP(hipGetLastError());
}

template <typename T, int MODE, int NB, int W>
static void launch_stencil_vb(const StencilParams &p, int zc, int order, const T *x, T *out, const T *v, const T *r,
                              const double *lam, hipStream_t s) {
    constexpr int VX = Vec16<T>::N;
    const int n = p.n;
    const long N = (long)n * n * n;
    if (zc <= 0 || zc > n) zc = n;
    const int nzc = (n + zc - 1) / zc;
    const unsigned grid = st_grid((n + 16 * VX - 1) / (16 * VX), (n + 15) / 16, nzc, (p.nb + NB - 1) / NB, order);
    hipLaunchKernelGGL((k_stencil_vb<T, MODE, NB, W>), dim3(grid), dim3(16, 16, 1), 0, s, x, out, v, r, lam, n, N, zc,
                       p.nb, order, (T)(1.0 / (p.h * p.h)), (T)p.scale, (T)p.omega, (T)p.diag_t,
                       static_cast<const T *>(p.x2), static_cast<T *>(p.out2), p.nb1 < 0 ? p.nb : p.nb1);
    LIRA_HIP(hipGetLastError());
}

// launch the variant's kernel for window W (4 bands per lane exist only at W = 9: their registers)
template <typename T, int W>
static void stencil_dispatch(int mode, const StVariant &sv, const StencilParams &p, const T *x, T *out, const T *v,
                             const T *r, const double *lam, hipStream_t s) {
#define LIRA_VB(MODE_, NB_) launch_stencil_vb<T, MODE_, NB_, W>(p, sv.zc, sv.order, x, out, v, r, lam, s)
#define LIRA_SV(MODE_, RPT_) launch_stencil_vec<T, MODE_, RPT_, W>(p, sv.zc, sv.order, x, out, v, r, lam, s)
#define LIRA_ST(MODE_, RPT_) launch_stencil<T, MODE_, RPT_, W, 64>(p, sv.zc, sv.order, x, out, v, r, lam, s)
#define LIRA_ST32(MODE_, RPT_) launch_stencil<T, MODE_, RPT_, W, 32>(p, sv.zc, sv.order, x, out, v, r, lam, s)
    if (sv.vec == 2) {
        bool four = false;
        if constexpr (W == 9) {
            if (sv.rpt == 4) {
                four = true;
                if (mode == ST_LAP) LIRA_VB(ST_LAP, 4);
                else if (mode == ST_HAM) LIRA_VB(ST_HAM, 4);
                else LIRA_VB(ST_JAC, 4);
            }
        }
        if (!four) {
            if (mode == ST_LAP) LIRA_VB(ST_LAP, 2);
            else if (mode == ST_HAM) LIRA_VB(ST_HAM, 2);
            else LIRA_VB(ST_JAC, 2);
        }
    } else if (sv.vec == 1) {
        if (sv.rpt == 1) {
            if (mode == ST_LAP) LIRA_SV(ST_LAP, 1);
            else if (mode == ST_HAM) LIRA_SV(ST_HAM, 1);
            else LIRA_SV(ST_JAC, 1);
        } else {
            if (mode == ST_LAP) LIRA_SV(ST_LAP, 2);
            else if (mode == ST_HAM) LIRA_SV(ST_HAM, 2);
            else LIRA_SV(ST_JAC, 2);
        }
    } else if (sv.vec == 3) {
        bool done = false;
        if constexpr (W == 9) {
            done = true;
            if (sv.rpt == 1) {
                if (mode == ST_LAP) LIRA_ST32(ST_LAP, 1);
                else if (mode == ST_HAM) LIRA_ST32(ST_HAM, 1);
                else LIRA_ST32(ST_JAC, 1);
            } else {
                if (mode == ST_LAP) LIRA_ST32(ST_LAP, 2);
                else if (mode == ST_HAM) LIRA_ST32(ST_HAM, 2);
                else LIRA_ST32(ST_JAC, 2);
            }
        }
        if (!done) throw InvalidError("stencil: 32-wide tiles exist only with the 9-plane window");
    } else if (sv.rpt == 4) {
        if (mode == ST_LAP) LIRA_ST(ST_LAP, 4);
        else if (mode == ST_HAM) LIRA_ST(ST_HAM, 4);
        else LIRA_ST(ST_JAC, 4);
    } else {
        if (mode == ST_LAP) LIRA_ST(ST_LAP, 2);
        else if (mode == ST_HAM) LIRA_ST(ST_HAM, 2);
        else LIRA_ST(ST_JAC, 2);
    }
#undef LIRA_VB
#undef LIRA_SV
#undef LIRA_ST
#undef LIRA_ST32
}

bool stencil_variant_usable(int variant, int n, int dtype_size) {
    if (variant < 0 || variant >= stencil_variants()) return false;
    const int vec = ST_VARIANTS[variant].vec;
    if (vec == 0 || vec == 3) return true;
    return n % (16 / dtype_size) == 0;
}

template <typename T>
void stencil(int mode, const StencilParams &p, const T *x, T *out, const T *v, const T *r, const double *lam,
             hipStream_t s, int variant) {
    if (p.nb <= 0) return;
    if (x == out) throw InvalidError("stencil: output aliases the input");
    if (p.nb1 >= 0 && p.nb1 < p.nb) {
        if (!p.x2 || !p.out2 || p.x2 == p.out2 || mode == ST_JAC)
            throw InvalidError("stencil: bad second band segment");
    }
    if (!stencil_buffer_ok(p.n, (int)sizeof(T)))
        throw InvalidError("stencil: grid too large for 32-bit field offsets ((n^3 + 4 n^2) values must stay below "
                           "2^31 bytes)");
    if (variant < 0) variant = g_default_variant[sizeof(T) == 8 ? 1 : 0];
    if (!stencil_variant_usable(variant, p.n, (int)sizeof(T))) variant = pick_variant(p.n, p.nb);
    if (ST_VARIANTS[variant].vec == 1 || ST_VARIANTS[variant].vec == 2) {
        // vector lanes need 16-byte aligned bases (every band then stays aligned since n is a multiple of the width)
        auto misaligned = [](const void *q) { return q && (reinterpret_cast<uintptr_t>(q) & 15) != 0; };
        if (misaligned(x) || misaligned(out) || misaligned(v) || misaligned(r) || misaligned(p.x2) ||
            misaligned(p.out2))
            variant = pick_variant(p.n, p.nb);
    }
    const StVariant &sv = ST_VARIANTS[variant];
    if (sv.win == 12) stencil_dispatch<T, 12>(mode, sv, p, x, out, v, r, lam, s);
    else stencil_dispatch<T, 9>(mode, sv, p, x, out, v, r, lam, s);
}


// ------------------------------------------------------------------------------------------ reductions ----
__device__ inline double block_sum256(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__device__ inline double block_max256(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

template <typename T>
__global__ __launch_bounds__(256) void k_reduce(RedJobs<T> jobs, long N, int n, double *partials) {
    __shared__ double sh[256];
    const int j = blockIdx.y;
    const int kind = jobs.kind[j];
    const T *a = jobs.a[j];
    const T *b = jobs.b[j];
    double acc = 0.0;
    const long stride = (long)gridDim.x * 256;
    if (kind == RED_DOT) {
        for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += stride) acc += (double)a[p] * (double)b[p];
    } else if (kind == RED_SUM) {
        for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += stride) acc += (double)a[p];
    } else if (kind == RED_ABSDIFF) {
        for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += stride) acc += fabs((double)a[p] - (double)b[p]);
    } else if (kind == RED_FACEMAX) {
        const long nn = (long)n * n;
        for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += stride) {
            const long k = p / nn, rem = p - k * nn, jj = rem / n, i = rem - jj * n;
            if (i == 0 || jj == 0 || k == 0 || i == n - 1 || jj == n - 1 || k == n - 1) acc = fmax(acc, fabs((double)a[p]));
        }
    } else {
        for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += stride) acc = fmax(acc, fabs((double)a[p]));
    }
    const bool is_max = kind == RED_FACEMAX || kind == RED_MAX_ABS;
    const double r = is_max ? block_max256(acc, sh) : block_sum256(acc, sh);
    if (threadIdx.x == 0) partials[(long)j * RED_BLOCKS + blockIdx.x] = r;
}

struct KindList {
    int kind[MAX_JOBS];
};

__global__ __launch_bounds__(256) void k_reduce_finish(KindList kinds, const double *partials, double *out,
                                                        double scale) {
    __shared__ double sh[256];
    const int j = blockIdx.x;
    const bool is_max = kinds.kind[j] == RED_FACEMAX || kinds.kind[j] == RED_MAX_ABS;
    double acc = 0.0;
    for (int p = threadIdx.x; p < RED_BLOCKS; p += 256) {
        const double v = partials[(long)j * RED_BLOCKS + p];
        acc = is_max ? fmax(acc, v) : acc + v;
    }
    const double r = is_max ? block_max256(acc, sh) : block_sum256(acc, sh);
    if (threadIdx.x == 0) out[j] = is_max ? r : r * scale;
}

template <typename T>
void reduce_jobs(const RedJobs<T> &jobs, int n, double scale, double *partials, double *out, hipStream_t s) {
    if (jobs.count <= 0) return;
    if (jobs.count > MAX_JOBS) throw InvalidError("reduce_jobs: too many jobs");
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_reduce<T>, dim3(RED_BLOCKS, jobs.count), dim3(256), 0, s, jobs, N, n, partials);
    LIRA_HIP(hipGetLastError());
    KindList kl;
    for (int j = 0; j < MAX_JOBS; j++) kl.kind[j] = j < jobs.count ? jobs.kind[j] : RED_SUM;
    hipLaunchKernelGGL(k_reduce_finish, dim3(jobs.count), dim3(256), 0, s, kl, partials, out, scale);
    LIRA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------- Gram ----
// G = scale S^T U over N points. A workgroup computes a 32 (S) x 64 (U) output block over one chunk of points,
// staging 64 points of its 96 columns in LDS at a time; each thread owns a 2 x 4 register tile accumulated in f64.

// A load through a pointer the compiler cannot place (one read from an LDS table of column bases) becomes a flat
// load, which also counts against the LDS wait counter: every wait for an LDS read would then wait for the loads in
// flight as well. Column data is always global memory, so these loads name that address space.
template <typename T> __device__ __forceinline__ T gload(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}

struct ColPtrs {
    const void *base[MAX_SEGS];
    int start[MAX_SEGS + 1];
    int nseg;
};

// column col's base address on the host (the same arithmetic as col_base)
static const void *col_host(const ColPtrs &c, int col, long N, int esize) {
    int s = 0;
    for (int i = 1; i < c.nseg; i++)
        if (col >= c.start[i]) s = i;
    return static_cast<const char *>(c.base[s]) + (size_t)(col - c.start[s]) * (size_t)N * (size_t)esize;
}

__device__ inline const void *col_base(const ColPtrs &c, int col, long N, int esize) {
    int s = 0;
#pragma unroll
    for (int i = 1; i < MAX_SEGS; i++)
        if (i < c.nseg && col >= c.start[i]) s = i;
    return static_cast<const char *>(c.base[s]) + (size_t)(col - c.start[s]) * (size_t)N * (size_t)esize;
}

// G = scale S^T U on the f64 matrix cores: v_mfma_f64_16x16x4f64 with A = S-columns x 4 points, B = 4 points x
// U-columns (lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; the 16 x 16 result lands as D[l / 16 + 4 r][l % 16]
// in register r, measured on gfx950). A workgroup (4 waves) owns a 32 x 64 output tile over one chunk of points; each wave two 16 x 16
// sub-tiles. Columns are staged 32 points at a time in LDS as f64 (f32 inputs convert exactly), so the products are
// exact and the accumulation is f64 in every precision mode.
typedef double lira_d4 __attribute__((ext_vector_type(4)));
constexpr int G2_TI = 32, G2_TJ = 64;

// Each wave streams its own 16-point stages through a private LDS slice (96 columns x 16 points): no block barrier
// in the loop, so the four waves of a workgroup never wait on each other. A stage is 24 elements per lane
// (consecutive lanes on consecutive points of one column: 16 lanes read 16 contiguous values), loaded into registers
// one stage ahead so their latency hides behind the 32 MFMAs of the current stage. The four waves take interleaved
// stages of the workgroup's chunk; their 32 x 64 accumulators are summed in a fixed order at the end.
constexpr int G2_S = 16;  // points per wave stage
// A wave's slab slice holds the 4 LOADS rows a stage loads (the tile's 32 S rows, then the packed U rows) and, last,
// a row that stays zero (the B operand of a U column past q): sized by LOADS, the slab of the common 16-load shape
// takes 35 KB instead of the 53 KB of the widest, and four workgroups fit a CU's LDS instead of two.
template <int LOADS> constexpr int g2_rows() { return 4 * LOADS + 1; }

// Every load is issued unconditionally and its value is not touched until the stage is staged: a missing column or a
// point past the chunk reads a zero constant instead. (A load under a branch, or a select on the loaded value, makes
// the compiler wait for the load at once, and the stage-ahead prefetch would never be in flight.)
template <typename T> struct GramZero;
template <> struct GramZero<float> { static __device__ const float *ptr(); };
template <> struct GramZero<double> { static __device__ const double *ptr(); };
__device__ const float lira_gram_zero_f = 0.0f;
__device__ const double lira_gram_zero_d = 0.0;
__device__ const float *GramZero<float>::ptr() { return &lira_gram_zero_f; }
__device__ const double *GramZero<double>::ptr() { return &lira_gram_zero_d; }

template <typename T, int LOADS>
__device__ inline void gram_fetch(const T *const *cols, long p0, long p_end, T (&v)[LOADS]) {
    const int lane = threadIdx.x & 63;
    const T *zero = GramZero<T>::ptr();
#pragma unroll
    for (int u = 0; u < LOADS; u++) {
        const int e = lane + 64 * u;
        const int col = e >> 4, pp = e & 15;
        const long p = p0 + pp;
        const T *c = cols[col];
        v[u] = gload(c != nullptr && p < p_end ? c + p : zero);
    }
}

// With sym (U's first m columns are S's), the 16 x 16 blocks of S^T S strictly below its diagonal are skipped
// (gram_mirrored): k_gram_finish reads their values from the transposed block, the same sum of the same exact products.
__host__ __device__ inline bool gram_mirrored(int sym, int m, int i, int j) {
    const int rb = i >> 4, cb = j >> 4;
    return sym && cb < rb && 16 * (cb + 1) <= m;
}

// The slab holds the tile's 32 S columns in rows 0..31 and, packed after them, only the U columns that are neither past
// q nor one of the tile's S columns (U = [S, R, HW, HP] in LOBPCG): each U column reads its B operand from its packed
// row, its S row, or the zero row. LOADS = 8 (the S rows) + a quarter of the packed U columns, rounded up to a
// multiple of 4: the elements each lane loads per 16-point stage.
// At least three waves per SIMD: the compiler then keeps the accumulators in VGPRs (128 registers at 16 loads, four
// waves) instead of 138 VGPRs + 64 AGPRs (two waves), with no spills; the loads of more waves are in flight at once.
template <typename T, int LOADS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_gram(ColPtrs S, ColPtrs U, int m,
                                                                                     int q, long N, long chunk,
                                              double *partials, int sym) {
    constexpr int ROWS = g2_rows<LOADS>(), ZERO_ROW = ROWS - 1;
    static_assert(4 * ROWS * (G2_S + 1) >= 3 * 16 * 64, "the final reduction's half pass must fit the slab");
    __shared__ double slab[4][ROWS][G2_S + 1];  // one slice per wave; reused for the final reduction
    __shared__ const T *cols[G2_TI + G2_TJ];
    __shared__ int urow[G2_TJ];  // the slab row each U column's B operand is read from
    const int nbj = (q + G2_TJ - 1) / G2_TJ;
    const int bi = blockIdx.y / nbj, bj = blockIdx.y - bi * nbj;
    const int i0 = bi * G2_TI, j0 = bj * G2_TJ;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    // which 16 x 16 output blocks this tile computes (uniform over the workgroup); a tile with none returns here,
    // before any barrier
    const int ni = min(2, (m - i0 + 15) / 16), nj = min(4, (q - j0 + 15) / 16);
    bool live[2][4];
    bool any = false;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            live[a][b] = a < ni && b < nj && !gram_mirrored(sym, m, i0 + 16 * a, j0 + 16 * b);
            any = any || live[a][b];
        }
    if (!any) return;
    if (t < G2_TI) cols[t] = i0 + t < m ? static_cast<const T *>(col_base(S, i0 + t, N, sizeof(T))) : nullptr;
    if (t >= 64 + G2_TI && t < 64 + G2_TI + G2_TJ) cols[t - 64] = nullptr;  // packed U rows default to zero loads
    __syncthreads();
    if (wave == 1) {
        // U column c = lane: an alias of one of the tile's S columns, a real column (packed in ballot order), or none
        const int c = lane;
        const T *uc = j0 + c < q ? static_cast<const T *>(col_base(U, j0 + c, N, sizeof(T))) : nullptr;
        int srow = -1;
        if (uc)
            for (int i = 0; i < G2_TI; i++)
                if (cols[i] == uc) {
                    srow = i;
                    break;
                }
        const bool real = uc != nullptr && srow < 0;
        const unsigned long long mask = __ballot(real);
        const int rank = __popcll(mask & ((1ull << c) - 1ull));
        urow[c] = srow >= 0 ? srow : real ? G2_TI + rank : ZERO_ROW;
        if (real) cols[G2_TI + rank] = uc;
    }
    // the zero row of every wave's slice
    if (lane < G2_S) slab[wave][ZERO_ROW][lane] = 0.0;
    lira_d4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = lira_d4{0.0, 0.0, 0.0, 0.0};
    const long p_begin = (long)blockIdx.x * chunk;
    const long p_end = min(N, p_begin + chunk);
    __syncthreads();
    double(*sl)[G2_S + 1] = slab[wave];
    // two stage buffers in registers: the loads of the stage after next are in flight while a stage is consumed
    T va[LOADS], vb[LOADS];
    long p0 = p_begin + (long)wave * G2_S;
    gram_fetch<T, LOADS>(cols, p0, p_end, va);
    gram_fetch<T, LOADS>(cols, p0 + 4 * G2_S, p_end, vb);
    const int r16 = lane & 15, k4 = lane >> 4;
    int brow[4];
#pragma unroll
    for (int b = 0; b < 4; b++) brow[b] = urow[16 * b + r16];
    // one stage: stage v into the wave's slice, refill v with the stage two ahead (past the chunk: zeros), then the
    // MFMAs of the staged points
    auto stage = [&](T(&v)[LOADS], long pn) {
#pragma unroll
        for (int u = 0; u < LOADS; u++) {
            const int e = lane + 64 * u;
            sl[e >> 4][e & 15] = (double)v[u];
        }
        __builtin_amdgcn_wave_barrier();
        gram_fetch<T, LOADS>(cols, pn, p_end, v);
#pragma unroll
        for (int kk = 0; kk < G2_S; kk += 4) {
            const double a0 = sl[r16][kk + k4];
            const double a1 = sl[16 + r16][kk + k4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                if (live[0][b] || live[1][b]) {
                    const double bv = sl[brow[b]][kk + k4];
                    if (live[0][b]) acc[0][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv, acc[0][b], 0, 0, 0);
                    if (live[1][b]) acc[1][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv, acc[1][b], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    };
    // stages in pairs (a stage past the chunk holds zeros and adds nothing)
    for (; p0 < p_end; p0 += 8 * G2_S) {
        stage(va, p0 + 8 * G2_S);
        stage(vb, p0 + 12 * G2_S);
    }
    // sum the four waves' accumulators in wave order through LDS, one 16 x 64 half (acc[a]) per pass: 16 doubles per
    // lane per wave
    double *red = &slab[0][0][0];
    double *dst = partials + (long)blockIdx.x * m * q;
    const int jl = lane & 15, ib = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; a++) {
        __syncthreads();
        if (wave > 0) {
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) red[(((wave - 1) * 16) + (b * 4 + r)) * 64 + lane] = acc[a][b][r];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int slot = b * 4 + r;
                    const double sum = ((acc[a][b][r] + red[(0 * 16 + slot) * 64 + lane]) +
                                        red[(1 * 16 + slot) * 64 + lane]) + red[(2 * 16 + slot) * 64 + lane];
                    const int i = i0 + a * 16 + ib + 4 * r, j = j0 + b * 16 + jl;
                    if (i < m && j < q) dst[(long)i * q + j] = sum;
                }
        }
    }
}

// sums the per-chunk partials of every output in a fixed order: 4 outputs per workgroup, 64 lanes per output each
// summing a strided subset of the chunks, then a fixed shuffle tree
__global__ __launch_bounds__(256) void k_gram_finish(const double *partials, int chunks, int m, int q, double scale,
                                                      double *G, int sym) {
    const int lane = threadIdx.x & 63;
    const long mq = (long)m * q;
    const long out = blockIdx.x * 4L + (threadIdx.x >> 6);
    // a skipped block's output is summed from its transposed entry (i, j) -> (j, i)
    long src = out;
    if (out < mq) {
        const int i = (int)(out / q), j = (int)(out - (long)i * q);
        if (gram_mirrored(sym, m, i, j)) src = (long)j * q + i;
    }
    double acc = 0.0;
    if (out < mq)
        for (int c = lane; c < chunks; c += 64) acc += partials[(long)c * mq + src];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0 && out < mq) G[out] = acc * scale;
}

// about 768 workgroups in all: three per CU in one round (the 16-load shape fits four per CU; 512 or 1024 measured
// slower at the LOBPCG shapes of a 95^3 grid)
static int gram_chunks(int m, int q, long N) {
    const int blocks = ((m + G2_TI - 1) / G2_TI) * ((q + G2_TJ - 1) / G2_TJ);
    int chunks = (768 + blocks - 1) / blocks;
    const long max_chunks = (N + 4 * G2_S - 1) / (4 * G2_S);
    if (chunks > max_chunks) chunks = (int)max_chunks;
    if (chunks < 1) chunks = 1;
    return chunks;
}

size_t gram_partials_needed(int m, int q, long N) { return (size_t)gram_chunks(m, q, N) * m * q; }

static ColPtrs to_ptrs(const Cols &c) {
    ColPtrs p{};
    p.nseg = c.nseg;
    int acc = 0;
    for (int i = 0; i < c.nseg; i++) {
        p.base[i] = c.seg[i].base;
        p.start[i] = acc;
        acc += c.seg[i].count;
    }
    p.start[c.nseg] = acc;
    return p;
}

template <typename T>
void gram(const Cols &S, const Cols &U, long N, double scale, double *partials, size_t partial_cap, double *G,
          hipStream_t s) {
    const int m = S.total(), q = U.total();
    if (m == 0 || q == 0) return;
    const int chunks = gram_chunks(m, q, N);
    if ((size_t)chunks * m * q > partial_cap) throw InvalidError("gram: partial buffer too small");
    long chunk = (N + chunks - 1) / chunks;
    chunk = ((chunk + 4 * G2_S - 1) / (4 * G2_S)) * (4 * G2_S);
    const int blocks = ((m + G2_TI - 1) / G2_TI) * ((q + G2_TJ - 1) / G2_TJ);
    // U begins with S's m columns (LOBPCG's U = [S, R, HW, HP]): S^T S is symmetric and its lower blocks are mirrored
    const ColPtrs sp = to_ptrs(S), up = to_ptrs(U);
    int sym = q >= m ? 1 : 0;
    for (int j = 0; j < m && sym; j++)
        if (col_host(sp, j, N, sizeof(T)) != col_host(up, j, N, sizeof(T))) sym = 0;
    // the packed U columns of the fullest tile: those neither past q nor among the tile's S columns
    int packed = 0;
    const int nbi = (m + G2_TI - 1) / G2_TI, nbj = (q + G2_TJ - 1) / G2_TJ;
    for (int bi = 0; bi < nbi; bi++) {
        std::vector<const void *> sc;
        for (int i = bi * G2_TI; i < std::min(m, (bi + 1) * G2_TI); i++) sc.push_back(col_host(sp, i, N, sizeof(T)));
        for (int bj = 0; bj < nbj; bj++) {
            int cnt = 0;
            for (int j = bj * G2_TJ; j < std::min(q, (bj + 1) * G2_TJ); j++)
                if (std::find(sc.begin(), sc.end(), col_host(up, j, N, sizeof(T))) == sc.end()) cnt++;
            packed = std::max(packed, cnt);
        }
    }
    const int loads = 8 + (packed + 15) / 16 * 4;  // 12, 16, 20 or 24 elements per lane per stage
    const dim3 grid(chunks, blocks);
    if (loads <= 12) hipLaunchKernelGGL((k_gram<T, 12>), grid, dim3(256), 0, s, sp, up, m, q, N, chunk, partials, sym);
    else if (loads <= 16) hipLaunchKernelGGL((k_gram<T, 16>), grid, dim3(256), 0, s, sp, up, m, q, N, chunk, partials, sym);
    else if (loads <= 20) hipLaunchKernelGGL((k_gram<T, 20>), grid, dim3(256), 0, s, sp, up, m, q, N, chunk, partials, sym);
    else hipLaunchKernelGGL((k_gram<T, 24>), grid, dim3(256), 0, s, sp, up, m, q, N, chunk, partials, sym);
    LIRA_HIP(hipGetLastError());
    const long mq = (long)m * q;
    hipLaunchKernelGGL(k_gram_finish, dim3((unsigned)((mq + 3) / 4)), dim3(256), 0, s, partials, chunks, m, q, scale, G,
                       sym);
    LIRA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------ lincomb ----
// Y_j = sum_i S_i C_ij: one thread per point holds QB outputs in f64 registers; grid.y covers q in groups of QB.
// C arrives padded (lincomb_rows(m) rows, lincomb_ldc(q) columns, zeros beyond m x q), so every loop below has a
// compile-time trip count and no bounds branch: each row's QB coefficients are wave-uniform and come in a few wide
// scalar loads, and every FMA takes one as its scalar operand. Padded rows read column 0 (their coefficients are 0).
constexpr int LC_MAX_M = 384;  // input columns of one lincomb: [X, W, P] of blocks up to 128 columns
constexpr int LC_GROUP = 8;    // input columns loaded together

static int lincomb_qb(int q) {
    // the output group QB: the smallest of 8, 16, 24, 32 that splits q into equal-count launches
    const int groups = (q + 31) / 32;
    const int per = (q + groups - 1) / groups;
    return per <= 8 ? 8 : per <= 16 ? 16 : per <= 24 ? 24 : 32;
}
int lincomb_rows(int m) { return (m + LC_GROUP - 1) / LC_GROUP * LC_GROUP; }
int lincomb_ldc(int q) {
    const int qb = lincomb_qb(q);
    return (q + qb - 1) / qb * qb;
}

template <typename T, int QB>
__global__ __launch_bounds__(256) void k_lincomb(ColPtrs S, int m, int mp, const double *__restrict__ C, int ldc,
                                                 int q, T *__restrict__ Y, int qsplit, T *__restrict__ Y2, long N) {
    __shared__ const T *cols[LC_MAX_M];
    const int j0 = blockIdx.y * QB;
    for (int i = threadIdx.x; i < mp; i += 256)
        cols[i] = static_cast<const T *>(col_base(S, i < m ? i : 0, N, sizeof(T)));
    __syncthreads();
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= N) return;
    double acc[QB];
#pragma unroll
    for (int c = 0; c < QB; c++) acc[c] = 0.0;
    for (int i0 = 0; i0 < mp; i0 += LC_GROUP) {
        // the group's loads are independent and issued together
        double sv[LC_GROUP];
#pragma unroll
        for (int u = 0; u < LC_GROUP; u++) sv[u] = (double)gload(cols[i0 + u] + p);
#pragma unroll
        for (int u = 0; u < LC_GROUP; u++) {
            const double *__restrict__ row = C + (long)(i0 + u) * ldc + j0;
#pragma unroll
            for (int c = 0; c < QB; c++) acc[c] += sv[u] * row[c];
        }
    }
#pragma unroll
    for (int c = 0; c < QB; c++) {
        const int j = j0 + c;
        if (j < q) {
            T *dst = j < qsplit ? Y + (long)j * N : Y2 + (long)(j - qsplit) * N;
            dst[p] = (T)acc[c];
        }
    }
}

template <typename T>
void lincomb(const Cols &S, const double *C, int ldc, int q, T *Y, long N, hipStream_t s, int qsplit, T *Y2) {
    const int m = S.total();
    if (q <= 0) return;
    if (m > LC_MAX_M) throw InvalidError("lincomb: more than " + std::to_string(LC_MAX_M) + " input columns");
    if (ldc != lincomb_ldc(q)) throw InvalidError("lincomb: coefficients not padded to lincomb_ldc(q) columns");
    if (qsplit < 0 || qsplit > q) qsplit = q;
    const unsigned gx = (unsigned)((N + 255) / 256);
    const int qb = lincomb_qb(q), mp = lincomb_rows(m);
    const ColPtrs sp = to_ptrs(S);
    const dim3 grid(gx, (unsigned)(ldc / qb));
#define LIRA_LC(QB_) hipLaunchKernelGGL((k_lincomb<T, QB_>), grid, dim3(256), 0, s, sp, m, mp, C, ldc, q, Y, qsplit, Y2, N)
    if (qb == 8) LIRA_LC(8);
    else if (qb == 16) LIRA_LC(16);
    else if (qb == 24) LIRA_LC(24);
    else LIRA_LC(32);
#undef LIRA_LC
    LIRA_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------------------- pointwise ----
#define LIRA_GRID1(N) dim3((unsigned)(((N) + 255) / 256)), dim3(256)

template <typename T> __global__ void k_fill(T *y, long n, T a) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < n) y[i] = a;
}
template <typename T> void fill(T *y, long count, double a, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, LIRA_GRID1(count), 0, s, y, count, (T)a);
    LIRA_HIP(hipGetLastError());
}

template <typename T> __global__ void k_axpy(T *y, const T *x, long n, T a) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
template <typename T> void axpy(T *y, const T *x, long count, double a, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy<T>, LIRA_GRID1(count), 0, s, y, x, count, (T)a);
    LIRA_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_residual(T *r, const T *hx, const T *x, const double *theta, int nb, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= N) return;
    const long id = (long)b * N + i;
    r[id] = (T)((double)hx[id] - theta[b] * (double)x[id]);
}
template <typename T> void residual(T *r, const T *hx, const T *x, const double *theta, int nb, long N, hipStream_t s) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_residual<T>, dim3((unsigned)((N + 255) / 256), nb), dim3(256), 0, s, r, hx, x, theta, nb, N);
    LIRA_HIP(hipGetLastError());
}

// R = HX - theta X and the preconditioner's first pass W = R / max(diag_t + V - lam, 0.1) in one pass (the same
// per-point formulas as k_residual and k_precond_init)
template <typename T>
__global__ void k_residual_precond(T *r, T *w, const T *hx, const T *x, const double *theta, const T *v,
                                   const double *lam, T diag_t, int nb, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= N) return;
    const long id = (long)b * N + i;
    const T rv = (T)((double)hx[id] - theta[b] * (double)x[id]);
    T d = diag_t + v[i] - (T)lam[b];
    d = d > T(0.1) ? d : T(0.1);
    r[id] = rv;
    w[id] = rv / d;
}
template <typename T>
void residual_precond(T *r, T *w, const T *hx, const T *x, const double *theta, const T *v, const double *lam,
                      double diag_t, int nb, long N, hipStream_t s) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_residual_precond<T>, dim3((unsigned)((N + 255) / 256), nb), dim3(256), 0, s, r, w, hx, x,
                       theta, v, lam, (T)diag_t, nb, N);
    LIRA_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_precond_init(T *w, const T *r, const T *v, const double *lam, T diag_t, int nb, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= N) return;
    T d = diag_t + v[i] - (T)lam[b];
    d = d > T(0.1) ? d : T(0.1);
    w[(long)b * N + i] = r[(long)b * N + i] / d;
}
template <typename T>
void precond_init(T *w, const T *r, const T *v, const double *lam, double diag_t, int nb, long N, hipStream_t s) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_precond_init<T>, dim3((unsigned)((N + 255) / 256), nb), dim3(256), 0, s, w, r, v, lam,
                       (T)diag_t, nb, N);
    LIRA_HIP(hipGetLastError());
}

template <typename T> __global__ void k_density(T *rho, const T *psi, const double *f, int nb, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    double acc = 0.0;
    for (int b = 0; b < nb; b++) {
        const double fb = f[b];
        if (fb == 0.0) continue;
        const double p = (double)psi[(long)b * N + i];
        acc += fb * p * p;
    }
    rho[i] = (T)acc;
}
template <typename T> void density(T *rho, const T *psi, const double *f, int nb, long N, hipStream_t s) {
    hipLaunchKernelGGL(k_density<T>, LIRA_GRID1(N), 0, s, rho, psi, f, nb, N);
    LIRA_HIP(hipGetLastError());
}

template <typename T> __global__ void k_sum3(T *o, const T *a, const T *b, const T *c, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < N) o[i] = (T)(((double)a[i] + (double)b[i]) + (double)c[i]);
}
template <typename T> void sum3(T *out, const T *a, const T *b, const T *c, long N, hipStream_t s) {
    hipLaunchKernelGGL(k_sum3<T>, LIRA_GRID1(N), 0, s, out, a, b, c, N);
    LIRA_HIP(hipGetLastError());
}
template <typename T> __global__ void k_add2(T *o, const T *a, const T *b, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < N) o[i] = a[i] + b[i];
}
template <typename T> void add2(T *out, const T *a, const T *b, long N, hipStream_t s) {
    hipLaunchKernelGGL(k_add2<T>, LIRA_GRID1(N), 0, s, out, a, b, N);
    LIRA_HIP(hipGetLastError());
}
template <typename T> __global__ void k_sub2(T *o, const T *a, const T *b, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < N) o[i] = a[i] - b[i];
}
template <typename T> void sub2(T *out, const T *a, const T *b, long N, hipStream_t s) {
    hipLaunchKernelGGL(k_sub2<T>, LIRA_GRID1(N), 0, s, out, a, b, N);
    LIRA_HIP(hipGetLastError());
}
template <typename T> void copy(T *out, const T *a, long N, hipStream_t s) {
    LIRA_HIP(hipMemcpyAsync(out, a, (size_t)N * sizeof(T), hipMemcpyDeviceToDevice, s));
}
template <typename Td, typename Ts> __global__ void k_convert(Td *o, const Ts *a, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < N) o[i] = (Td)a[i];
}
template <typename Td, typename Ts> void convert(Td *out, const Ts *a, long N, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<Td, Ts>), LIRA_GRID1(N), 0, s, out, a, N);
    LIRA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------- LSDA ----
// Dirac exchange + PW92 correlation (qm/src/xc.rs:78-187), evaluated in f64 at every point.
template <typename T> __global__ void k_xc(const T *nu, const T *nd, T *e, T *vu, T *vd, long N, const T *pc) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    double ee, a, b;
    // with a partial core the LSDA sees n_sigma = rho_sigma + rho_pc / 2 (qm/src/scf.rs:757-769)
    const double c = pc ? 0.5 * (double)pc[i] : 0.0;
    if (pc) exc_point((double)nu[i] + c, (double)nd[i] + c, ee, a, b);
    else exc_point((double)nu[i], (double)nd[i], ee, a, b);
    e[i] = (T)ee;
    vu[i] = (T)a;
    vd[i] = (T)b;
}
template <typename T> void xc_lsda(const T *nu, const T *nd, T *e, T *vu, T *vd, long N, hipStream_t s, const T *pc) {
    hipLaunchKernelGGL(k_xc<T>, LIRA_GRID1(N), 0, s, nu, nd, e, vu, vd, N, pc);
    LIRA_HIP(hipGetLastError());
}

// rho_in = max(2 rho_t - rho_{t-1}, 0): the density extrapolation of a warm start (crates/qm-gpu/src/scf.rs:1146-1162)
template <typename T> __global__ void k_extrapolate(T *out, const T *a, const T *b, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const double v = 2.0 * (double)a[i] - (double)b[i];
    out[i] = (T)(v > 0.0 ? v : 0.0);
}
template <typename T> void extrapolate(T *out, const T *a, const T *b, long N, hipStream_t s) {
    hipLaunchKernelGGL(k_extrapolate<T>, LIRA_GRID1(N), 0, s, out, a, b, N);
    LIRA_HIP(hipGetLastError());
}

// y *= num / den when both are > 0 (the per-channel charge renormalisation of the mixed density to the output's
// integral, crates/qm-gpu/src/scf.rs:1574-1582): the ratio in f64, rounded to T, applied in T
template <typename T> __global__ void k_scale_ratio(T *y, long N, const double *num, const double *den) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const double a = *num, b = *den;
    if (!(a > 0.0 && b > 0.0)) return;
    y[i] = y[i] * (T)(a / b);
}
template <typename T> void scale_ratio(T *y, long N, const double *num, const double *den, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_ratio<T>, LIRA_GRID1(N), 0, s, y, N, num, den);
    LIRA_HIP(hipGetLastError());
}

// (u + d)(p) by trilinear interpolation: the cell and fractions of qm::Grid::trilinear_cell in f64, neighbours beyond an
// upper face read 0, blended a (1 - t) + b t along x, then y, then z (crates/qm/src/grid.rs:54-86)
template <typename T>
__global__ void k_sample(const T *u, const T *d, int n, double h, const double *pts, int count, double *out) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= count) return;
    const double o = -((double)n - 1.0) * h / 2.0;
    long cell[3];
    double t[3];
    for (int a = 0; a < 3; a++) {
        const double x = (pts[3 * q + a] - o) / h;
        if (!(x >= 0.0 && x <= (double)(n - 1))) {
            out[q] = 0.0;
            return;
        }
        const double i0 = floor(x);
        cell[a] = (long)i0;
        t[a] = x - i0;
    }
    auto at = [&](int di, int dj, int dk) {
        const long a = cell[0] + di, b = cell[1] + dj, c = cell[2] + dk;
        if (a < n && b < n && c < n) {
            const long id = a + (long)n * (b + (long)n * c);
            return (double)u[id] + (double)d[id];
        }
        return 0.0;
    };
    auto lerp = [](double a, double b, double tt) { return a * (1.0 - tt) + b * tt; };
    auto xl = [&](int dj, int dk) { return lerp(at(0, dj, dk), at(1, dj, dk), t[0]); };
    auto yl = [&](int dk) { return lerp(xl(0, dk), xl(1, dk), t[1]); };
    out[q] = lerp(yl(0), yl(1), t[2]);
}
template <typename T>
void sample_density(const T *u, const T *d, int n, double h, const double *pts, int count, double *out,
                    hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_sample<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, u, d, n, h, pts, count, out);
    LIRA_HIP(hipGetLastError());
}

// --------------------------------------------------------------------------------------------- nuclei ----
template <typename T>
__global__ void k_nuclear(T *out, int n, double h, const double *pos, const double *zf, int nat, double sigma) {
    const long N = (long)n * n * n;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const long nn = (long)n * n;
    const long k = i / nn, rem = i - k * nn, j = rem / n, ii = rem - j * n;
    const double o = -((double)n - 1.0) * h / 2.0;
    const double x = o + (double)ii * h, y = o + (double)j * h, z = o + (double)k * h;
    const double s2 = sigma * sqrt(2.0), on_site = sqrt(2.0 / M_PI) / sigma;
    double acc = 0.0;
    for (int a = 0; a < nat; a++) {
        const double dx = x - pos[3 * a], dy = y - pos[3 * a + 1], dz = z - pos[3 * a + 2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        acc += d < 1e-10 ? -zf[a] * on_site : -zf[a] * erf(d / s2) / d;
    }
    out[i] = (T)acc;
}
template <typename T>
void nuclear_potential(T *out, int n, double h, const double *pos, const double *zf, int nat, double sigma,
                       hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_nuclear<T>, LIRA_GRID1(N), 0, s, out, n, h, pos, zf, nat, sigma);
    LIRA_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_deposit(T *out, int n, double h, const double *pos, const double *zf, int nat, double sigma,
                          double scale) {
    const long N = (long)n * n * n;
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    const long nn = (long)n * n;
    const long k = i / nn, rem = i - k * nn, j = rem / n, ii = rem - j * n;
    const double o = -((double)n - 1.0) * h / 2.0;
    const double x = o + (double)ii * h, y = o + (double)j * h, z = o + (double)k * h;
    const double norm = 1.0 / pow(2.0 * M_PI * sigma * sigma, 1.5), inv = 1.0 / (2.0 * sigma * sigma);
    double acc = 0.0;
    for (int a = 0; a < nat; a++) {
        const double dx = x - pos[3 * a], dy = y - pos[3 * a + 1], dz = z - pos[3 * a + 2];
        acc += zf[a] * norm * exp(-(dx * dx + dy * dy + dz * dz) * inv);
    }
    out[i] = (T)(acc * scale);
}
template <typename T>
void deposit_charge(T *out, int n, double h, const double *pos, const double *zf, int nat, double sigma, double scale,
                    hipStream_t s) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_deposit<T>, LIRA_GRID1(N), 0, s, out, n, h, pos, zf, nat, sigma, scale);
    LIRA_HIP(hipGetLastError());
}

// The LCG x -> a x + c (mod 2^64) advanced t steps: (A_t, C_t) with x_t = A_t x_0 + C_t, by binary powering.
__device__ inline void lcg_jump(uint64_t t, uint64_t &A, uint64_t &Cc) {
    uint64_t a = 6364136223846793005ULL, c = 1442695040888963407ULL;
    A = 1;
    Cc = 0;
    while (t) {
        if (t & 1) {
            A = A * a;
            Cc = Cc * a + c;
        }
        c = c * a + c;
        a = a * a;
        t >>= 1;
    }
}

constexpr int IO_RUN = 64;
template <typename T>
__global__ void k_initial(T *out, int n, double h, const double *pos, int nat, int k, uint64_t seed0) {
    const long N = (long)n * n * n;
    const long total = N * k;
    const long e0 = (blockIdx.x * 256L + threadIdx.x) * IO_RUN;
    if (e0 >= total) return;
    uint64_t A, Cc;
    lcg_jump((uint64_t)e0, A, Cc);
    uint64_t seed = A * seed0 + Cc;
    const double o = -((double)n - 1.0) * h / 2.0;
    const long nn = (long)n * n;
    for (long e = e0; e < min(total, e0 + IO_RUN); e++) {
        seed = seed * 6364136223846793005ULL + 1442695040888963407ULL;
        const double u = (double)(seed >> 11) / 9007199254740992.0;
        const long i = e % N;
        const long kk = i / nn, rem = i - kk * nn, j = rem / n, ii = rem - j * n;
        const double x = o + (double)ii * h, y = o + (double)j * h, z = o + (double)kk * h;
        double env = 0.0;
        for (int a = 0; a < nat; a++) {
            const double dx = x - pos[3 * a], dy = y - pos[3 * a + 1], dz = z - pos[3 * a + 2];
            env += exp(-(dx * dx + dy * dy + dz * dz) / 4.0);
        }
        out[e] = (T)((u + 0.5) * env);
    }
}
template <typename T>
void initial_orbitals(T *out, int n, double h, const double *pos, int nat, int k, uint64_t seed0, hipStream_t s) {
    const long total = (long)n * n * n * k;
    const long threads = (total + IO_RUN - 1) / IO_RUN;
    hipLaunchKernelGGL(k_initial<T>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, out, n, h, pos, nat, k,
                       seed0);
    LIRA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------- DIIS ----
constexpr int DIIS_MAX = 16;
template <typename T> struct PtrList {
    const T *p[DIIS_MAX];
};
template <typename T>
__global__ void k_diis_mix(T *out, PtrList<T> rho, PtrList<T> res, const double *c, int m, double alpha, long N) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= N) return;
    double acc = 0.0;
    for (int q = 0; q < m; q++) acc += c[q] * ((double)rho.p[q][i] + alpha * (double)res.p[q][i]);
    out[i] = (T)(acc > 0.0 ? acc : 0.0);
}
template <typename T>
void diis_mix(T *out, const T *const *rho, const T *const *res, const double *c, int m, double alpha, long N,
              hipStream_t s) {
    if (m > DIIS_MAX) throw InvalidError("diis_mix: history too long");
    PtrList<T> a{}, b{};
    for (int q = 0; q < m; q++) {
        a.p[q] = rho[q];
        b.p[q] = res[q];
    }
    hipLaunchKernelGGL(k_diis_mix<T>, LIRA_GRID1(N), 0, s, out, a, b, c, m, alpha, N);
    LIRA_HIP(hipGetLastError());
}

// -------------------------------------------------------------------------------------------- forces ----
template <typename T>
__global__ __launch_bounds__(256) void k_hf_forces(const T *rho, int n, double h, const double *pos, const double *zf,
                                                   double sigma, double *partials) {
    __shared__ double sh[256];
    const int I = blockIdx.y;
    const long N = (long)n * n * n, nn = (long)n * n;
    const double o = -((double)n - 1.0) * h / 2.0;
    const double a = sigma * sqrt(2.0);
    const double px = pos[3 * I], py = pos[3 * I + 1], pz = pos[3 * I + 2], z = zf[I];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += (long)gridDim.x * 256) {
        const long k = p / nn, rem = p - k * nn, j = rem / n, i = rem - j * n;
        const double dx = o + (double)i * h - px, dy = o + (double)j * h - py, dz = o + (double)k * h - pz;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (d < 1e-8) continue;
        const double u = d / a;
        const double gp = ((2.0 / sqrt(M_PI)) * u * exp(-u * u) - erf(u)) / (d * d);
        const double w = (double)rho[p] * (-z * gp / d);
        f0 += w * dx;
        f1 += w * dy;
        f2 += w * dz;
    }
    const double r0 = block_sum256(f0, sh), r1 = block_sum256(f1, sh), r2 = block_sum256(f2, sh);
    if (threadIdx.x == 0) {
        partials[((long)I * 3 + 0) * RED_BLOCKS + blockIdx.x] = r0;
        partials[((long)I * 3 + 1) * RED_BLOCKS + blockIdx.x] = r1;
        partials[((long)I * 3 + 2) * RED_BLOCKS + blockIdx.x] = r2;
    }
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *partials, double *out, double scale) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int p = threadIdx.x; p < RED_BLOCKS; p += 256) acc += partials[(long)blockIdx.x * RED_BLOCKS + p];
    const double r = block_sum256(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = r * scale;
}

template <typename T>
void hf_forces(const T *rho, int n, double h, const double *pos, const double *zf, int nat, double sigma,
               double *partials, double *out, hipStream_t s) {
    if (nat <= 0) return;
    hipLaunchKernelGGL(k_hf_forces<T>, dim3(RED_BLOCKS, nat), dim3(256), 0, s, rho, n, h, pos, zf, sigma, partials);
    LIRA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(3 * nat), dim3(256), 0, s, partials, out, h * h * h);
    LIRA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------- explicit instances ----
#define LIRA_INST(T)                                                                                                  \
    template void stencil<T>(int, const StencilParams &, const T *, T *, const T *, const T *, const double *,       \
                             hipStream_t, int);                                                                       \
    template void reduce_jobs<T>(const RedJobs<T> &, int, double, double *, double *, hipStream_t);                 \
    template void gram<T>(const Cols &, const Cols &, long, double, double *, size_t, double *, hipStream_t);       \
    template void residual_precond<T>(T *, T *, const T *, const T *, const double *, const T *, const double *,     \
                                      double, int, long, hipStream_t);                                         \
    template void lincomb<T>(const Cols &, const double *, int, int, T *, long, hipStream_t, int, T *);                       \
    template void fill<T>(T *, long, double, hipStream_t);                                                           \
    template void axpy<T>(T *, const T *, long, double, hipStream_t);                                                \
    template void residual<T>(T *, const T *, const T *, const double *, int, long, hipStream_t);                    \
    template void precond_init<T>(T *, const T *, const T *, const double *, double, int, long, hipStream_t);        \
    template void density<T>(T *, const T *, const double *, int, long, hipStream_t);                                \
    template void sum3<T>(T *, const T *, const T *, const T *, long, hipStream_t);                                  \
    template void add2<T>(T *, const T *, const T *, long, hipStream_t);                                             \
    template void sub2<T>(T *, const T *, const T *, long, hipStream_t);                                             \
    template void copy<T>(T *, const T *, long, hipStream_t);                                                        \
    template void xc_lsda<T>(const T *, const T *, T *, T *, T *, long, hipStream_t, const T *);                                \
    template void nuclear_potential<T>(T *, int, double, const double *, const double *, int, double, hipStream_t); \
    template void deposit_charge<T>(T *, int, double, const double *, const double *, int, double, double,          \
                                    hipStream_t);                                                                     \
    template void initial_orbitals<T>(T *, int, double, const double *, int, int, uint64_t, hipStream_t);            \
    template void diis_mix<T>(T *, const T *const *, const T *const *, const double *, int, double, long,            \
                              hipStream_t);                                                                           \
    template void hf_forces<T>(const T *, int, double, const double *, const double *, int, double, double *,        \
                               double *, hipStream_t);                                                                \
    template void extrapolate<T>(T *, const T *, const T *, long, hipStream_t);                                      \
    template void scale_ratio<T>(T *, long, const double *, const double *, hipStream_t);                             \
    template void sample_density<T>(const T *, const T *, int, double, const double *, int, double *, hipStream_t);
LIRA_INST(float)
LIRA_INST(double)
template void convert<float, double>(float *, const double *, long, hipStream_t);
template void convert<double, float>(double *, const float *, long, hipStream_t);
template void convert<double, double>(double *, const double *, long, hipStream_t);
template void convert<float, float>(float *, const float *, long, hipStream_t);

}  // namespace lira
