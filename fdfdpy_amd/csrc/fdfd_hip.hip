// fdfd_hip.hip -- gfx950 kernels and C ABI for the 2D FDFD solve path.
//
// Operator: the SC-PML 5-point stencil of fdfdpy/linalg.py:41-107 stored as
// five complex planes. Factor: a block sweep over x (blocks are y-columns of
// size n = ny, padded to TB) of the x-non-periodic part T of A:
//     S_0 = T_0,  S_k = T_k - diag(l_k) G_{k-1} diag(u_{k-1}),  G_k = S_k^-1
// with l_k / u_k the couplings to the previous / next column. Each G_k is
// formed by tiled Gauss-Jordan (TB x TB diagonal pivot tiles, inverted with
// 2 x 2 block pivots inside the tile, or with partial pivoting inside the
// tile in the fallback factor; rank-TB updates on MFMA). The periodic x-wrap
// of A (derivatives.py:22,25) is handled by a Woodbury correction formed
// after the block factorization; refinement against the exact operator
// removes what a truncated correction leaves.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>
#include <atomic>
#include <mutex>
#include <condition_variable>
#include <thread>
#include <deque>
#include <map>
#include <memory>

#include "../../include/fdfd_hip.h"

namespace {

constexpr int TB = 32;        // pivot tile / output tile edge
constexpr int TBP = TB + 1;   // padded LDS row
constexpr int NT = 256;       // threads per workgroup (4 waves)
constexpr int MAXJ = 32;      // factorization chains per batched step launch
constexpr int MAX_BATCH = 16; // operators per fdfd_factor_batch call (a twisted one is two chains)
constexpr int MAXR = 4;       // right-hand sides per sweep pass (per-block sweeps)
constexpr int MAX_BLOCK_N = 2048;   // largest block (the shorter grid side)
constexpr int NORM_PARTS = 512;

// constants.py:3-6
constexpr double EPSILON_0 = 8.85418782e-12;
constexpr double MU_0 = 1.25663706e-6;

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(FDFD_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
  } while (0)

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline int pad_n(int n) { return (n + TB - 1) / TB * TB; }

// ---------------------------------------------------------------- complex
struct cd { double r, i; };
__host__ __device__ inline cd cmul(cd a, cd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__host__ __device__ inline cd cadd(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }
__host__ __device__ inline cd csub(cd a, cd b) { return {a.r - b.r, a.i - b.i}; }
__host__ __device__ inline cd cscale(cd a, double s) { return {a.r * s, a.i * s}; }
// 1/z by Smith's method (no overflow of |z|^2)
template <typename T>
__host__ __device__ inline void crecip(T zr, T zi, T &outr, T &outi) {
  if (fabs(zr) >= fabs(zi)) {
    T t = zi / zr, den = zr + zi * t;
    outr = T(1) / den;
    outi = -t / den;
  } else {
    T t = zr / zi, den = zi + zr * t;
    outr = t / den;
    outi = T(-1) / den;
  }
}
__host__ __device__ inline cd cinv(cd z) { cd o; crecip<double>(z.r, z.i, o.r, o.i); return o; }

// ------------------------------------------------------- PML (pml.py:7-39)
// s = 1 - i sigma(l)/(omega eps0 L0), sigma = sigma_max (l/dw)^4,
// sigma_max = -(m+1) lnR / (2 eta0 dw), m = 4, lnR = -12.
__host__ __device__ inline cd s_of_l(double l, double dw, double omega, double L0) {
  const double eta0 = sqrt(MU_0 / EPSILON_0);
  const double sig_max = -(4.0 + 1.0) * (-12.0) / (2.0 * eta0 * dw);
  double q = l / dw;
  double sig = sig_max * (q * q) * (q * q);
  return {1.0, -sig / (omega * EPSILON_0 * L0)};
}
// kind 0 = 'f', 1 = 'b'; left layer i <= npml, right layer i > nw - npml.
__host__ __device__ inline cd sfactor(int kind, int i, int nw, int npml, double hw, double omega, double L0) {
  if (npml < 1) return {1.0, 0.0};
  double dw = npml * hw;
  if (i <= npml) {
    double l = kind == 0 ? hw * (npml - i + 0.5) : hw * (double)(npml - i + 1);
    return s_of_l(l, dw, omega, L0);
  }
  if (i > nw - npml) {
    double l = kind == 0 ? hw * (i - (nw - npml) - 0.5) : hw * (double)(i - (nw - npml) - 1);
    return s_of_l(l, dw, omega, L0);
  }
  return {1.0, 0.0};
}

__device__ inline cd ld_c(const double *p, size_t idx) { return {p[2 * idx], p[2 * idx + 1]}; }
__device__ inline void st_c(double *p, size_t idx, cd v) { p[2 * idx] = v.r; p[2 * idx + 1] = v.i; }

// ------------------------------------------ assembly (linalg.py:41-107)
__global__ void __launch_bounds__(NT) k_assemble(fdfd_grid g, const double *__restrict__ eps,
                                                 double *__restrict__ op) {
  const int j = blockIdx.x * NT + threadIdx.x;
  const int i = blockIdx.y;
  const int nx = g.nx, ny = g.ny;
  if (j >= ny) return;
  // cell sizes as dL() forms them: (N*dl)/N (linalg.py:25-33, simulation.py:33-34)
  const double hx = ((double)nx * g.dl) / nx, hy = ((double)ny * g.dl) / ny;
  const int ip = (i + 1) % nx, im = (i + nx - 1) % nx;
  const int jp = (j + 1) % ny, jm = (j + ny - 1) % ny;
  cd isxf = cinv(sfactor(0, i, nx, g.npml_x, hx, g.omega, g.L0));
  cd isxb = cinv(sfactor(1, i, nx, g.npml_x, hx, g.omega, g.L0));
  cd isxb_p = cinv(sfactor(1, ip, nx, g.npml_x, hx, g.omega, g.L0));
  cd isyf = cinv(sfactor(0, j, ny, g.npml_y, hy, g.omega, g.L0));
  cd isyb = cinv(sfactor(1, j, ny, g.npml_y, hy, g.omega, g.L0));
  cd isyb_p = cinv(sfactor(1, jp, ny, g.npml_y, hy, g.omega, g.L0));
  const double eps0_ = EPSILON_0 * g.L0, mu0_ = MU_0 * g.L0;
  const size_t M = (size_t)nx * ny, p = (size_t)i * ny + j;
  // stretched difference rows: Dxf row i = (1/sxf_i)/hx * (x[i+1]-x[i]),
  // Dxb row i = (1/sxb_i)/hx * (x[i]-x[i-1])
  cd fx = cscale(isxf, 1.0 / hx), fy = cscale(isyf, 1.0 / hy);
  cd bx = cscale(isxb, 1.0 / hx), bxp = cscale(isxb_p, 1.0 / hx);
  cd by = cscale(isyb, 1.0 / hy), byp = cscale(isyb_p, 1.0 / hy);
  cd xm, xp, ym, yp, cc;
  cd e = ld_c(eps, p);
  if (g.pol == FDFD_POL_EZ) {
    fx = cscale(fx, 1.0 / mu0_);
    fy = cscale(fy, 1.0 / mu0_);
    xp = cmul(fx, bxp);
    xm = cmul(fx, bx);
    yp = cmul(fy, byp);
    ym = cmul(fy, by);
    cd w2e = cscale(cscale(e, eps0_), g.omega * g.omega);
    cc = cadd(csub(csub(csub(csub({0, 0}, xp), xm), yp), ym), w2e);
  } else {
    // grid_average(eps0' eps, 'x') at (i,j) and (i+1,j); 'y' at (i,j), (i,j+1)
    cd e_im = ld_c(eps, (size_t)im * ny + j), e_ip = ld_c(eps, (size_t)ip * ny + j);
    cd e_jm = ld_c(eps, (size_t)i * ny + jm), e_jp = ld_c(eps, (size_t)i * ny + jp);
    cd ee = cscale(e, eps0_);
    cd ex_i = cscale(cadd(cscale(e_im, eps0_), ee), 0.5);
    cd ex_ip = cscale(cadd(ee, cscale(e_ip, eps0_)), 0.5);
    cd ey_j = cscale(cadd(cscale(e_jm, eps0_), ee), 0.5);
    cd ey_jp = cscale(cadd(ee, cscale(e_jp, eps0_)), 0.5);
    xp = cmul(fx, cmul(cinv(ex_ip), bxp));
    xm = cmul(fx, cmul(cinv(ex_i), bx));
    yp = cmul(fy, cmul(cinv(ey_jp), byp));
    ym = cmul(fy, cmul(cinv(ey_j), by));
    cc = cadd(csub(csub(csub(csub({0, 0}, xp), xm), yp), ym), {g.omega * g.omega * mu0_, 0.0});
  }
  st_c(op, FDFD_OP_C * M + p, cc);
  st_c(op, FDFD_OP_XM * M + p, xm);
  st_c(op, FDFD_OP_XP * M + p, xp);
  st_c(op, FDFD_OP_YM * M + p, ym);
  st_c(op, FDFD_OP_YP * M + p, yp);
}

// ------------------------------------------------- stencil y = A x / b - A x
// One thread per cell for all right-hand sides: the five coefficients are read
// once and applied to every x[r]. b may alias y (each element is read by the
// thread that writes it); x must not.
__global__ void __launch_bounds__(NT) k_stencil(int nx, int ny, int nrhs, const double *__restrict__ op,
                                                const double *__restrict__ x,
                                                const double *b, double *y) {
  const int j = blockIdx.x * NT + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= ny) return;
  const size_t M = (size_t)nx * ny, p = (size_t)i * ny + j;
  const int ip = (i + 1) % nx, im = (i + nx - 1) % nx;
  const int jp = (j + 1) % ny, jm = (j + ny - 1) % ny;
  const size_t pxm = (size_t)im * ny + j, pxp = (size_t)ip * ny + j;
  const size_t pym = (size_t)i * ny + jm, pyp = (size_t)i * ny + jp;
  const cd c0 = ld_c(op, FDFD_OP_C * M + p), cxm = ld_c(op, FDFD_OP_XM * M + p);
  const cd cxp = ld_c(op, FDFD_OP_XP * M + p), cym = ld_c(op, FDFD_OP_YM * M + p);
  const cd cyp = ld_c(op, FDFD_OP_YP * M + p);
  for (int r = 0; r < nrhs; ++r) {
    const double *xr = x + 2 * M * r;
    cd acc = cmul(c0, ld_c(xr, p));
    acc = cadd(acc, cmul(cxm, ld_c(xr, pxm)));
    acc = cadd(acc, cmul(cxp, ld_c(xr, pxp)));
    acc = cadd(acc, cmul(cym, ld_c(xr, pym)));
    acc = cadd(acc, cmul(cyp, ld_c(xr, pyp)));
    if (b) acc = csub(ld_c(b + 2 * M * r, p), acc);
    st_c(y + 2 * M * r, p, acc);
  }
}

// ------------------------------------------------------------------ MFMA
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator t in lane l (column is l % 16)
  static __device__ inline int crow(int l, int t) { return (l >> 4) + 4 * t; }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int crow(int l, int t) { return 4 * (l >> 4) + t; }
};

// Layout of the 16x16x4 instructions (wave64): lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; accumulator t of lane l is C[crow(l, t)][l%16], with
// crow = l/16 + 4t for f64 and 4(l/16) + t for f32.
__global__ void k_selftest_mfma_f64(const double *a, const double *b, double *d, int raw) {
  const int l = threadIdx.x;
  Mfma<double>::v4 c = {0, 0, 0, 0};
  c = Mfma<double>::mac(a[(l % 16) * 4 + l / 16], b[(l / 16) * 16 + l % 16], c);
  for (int t = 0; t < 4; ++t) {
    if (raw) d[l * 4 + t] = c[t];
    else d[Mfma<double>::crow(l, t) * 16 + l % 16] = c[t];
  }
}
__global__ void k_selftest_mfma_f32(const double *a, const double *b, double *d) {
  const int l = threadIdx.x;
  Mfma<float>::v4 c = {0, 0, 0, 0};
  c = Mfma<float>::mac((float)a[(l % 16) * 4 + l / 16], (float)b[(l / 16) * 16 + l % 16], c);
  for (int t = 0; t < 4; ++t) d[Mfma<float>::crow(l, t) * 16 + l % 16] = c[t];
}

// ------------------------------------------------ cross-workgroup exchange
// Data passed between workgroups of one launch is written and read with
// agent-scope relaxed atomics: coherent across the XCDs' L2 caches by
// themselves, so a flag needs only workgroup-scope ordering (the barrier's
// wait for every wave's stores) and no whole-L2 writeback or invalidate.
// Every wait is bounded: after SPIN_CAP polls it raises an abort flag, and
// every later wait returns at once, so a grid always drains.
constexpr unsigned SPIN_CAP = 1u << 21;
// host settings, read by the calling threads and the engines' scheduler threads
std::atomic<unsigned> g_spin_cap{SPIN_CAP};  // passed to the persistent kernels
std::atomic<int> g_step_sym{1};              // launch-per-step factorizations use symmetric steps
std::atomic<int> g_twist{1};                 // truncated-wrap factorizations are twisted where they fit

template <typename T>
__device__ inline T ld_xcd(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline void st_xcd(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Every wave waits for its own outstanding memory operations (the exchange
// stores among them); followed by a barrier, the data of the whole group is
// then visible before thread 0 raises a flag. (The barrier alone only waits
// for LDS traffic.)
__device__ inline void stores_complete() { __builtin_amdgcn_s_waitcnt(0); }

// Wave 0 waits until flags[0..nflags) are all >= target (or abort); then
// the whole group. Per-workgroup flags are plain stores: no same-address
// atomics to serialize at the coherence point.
__device__ inline void wait_all_flags(const unsigned *flags, int nflags, unsigned target, unsigned *abort,
                                      unsigned cap) {
  if (threadIdx.x < 64) {
    unsigned spins = 0;
    for (;;) {
      int ok = 1;
      for (int w = threadIdx.x; w < nflags; w += 64)
        ok &= __hip_atomic_load(flags + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target;
      if (__all(ok)) break;
      if ((++spins & 63) == 0 && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (spins > cap) {
        if (threadIdx.x == 0) __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// thread 0 waits until *flag >= target (or abort); then the whole group
__device__ inline void wait_flag(const unsigned *flag, unsigned target, unsigned *abort, unsigned cap) {
  if (threadIdx.x == 0) {
    unsigned spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if ((++spins & 63) == 0 && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (spins > cap) {
        __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// --------------------------------------------------------- factor: step job
template <typename T> struct FJob {
  const double *op;  // band storage of this operator
  T *G;              // [m][2][np][np] planar (re plane, im plane) inverses
  T *W;              // [2][np][np] ping-pong partner of G_k
  T *Pinv;           // [2 slots][2][TB][TB]
  int *status;       // set to 1 on a non-finite / zero pivot, 2 on a timed-out wait
  const T *dense;    // one dense matrix ([2][np][np]) to invert (m = 1); null: the operator's S_k
  T *rbuf;           // [nt][2][TB][TB] R = P X[p,J] of the step, by column J (symmetric jobs)
  unsigned *rflag;   // [nt] the chain step + 1 whose R[J] is in rbuf; null: no column workgroups
  T *dsym;           // [4][np] D, 1/D (re, im planes) of the block symmetrizer, written
                     // by the initial-pivot launch; null: every tile of every step
  int m, n, np, nt;  // blocks, block size, padded block size, tiles per edge
  int k, p;          // step performed by this launch; k < 0: initial pivot
  int dir;           // 1: blocks upwards from 0 (the whole chain, or the top chain of a twisted
                     // factorization, ending with its twist block c); -1: the bottom chain m-1..c+1
  int c;             // the twist block (m - 1: not twisted)
};
// Workgroup b < njobs takes the critical tile of job b (the tile inverted for
// the next step), so those start first; the rest of the first crow * XCDS
// workgroups idle. Then come the column workgroups of symmetric jobs (job j's
// are [rfirst[j], rfirst[j + 1]) after XCDS * crow): one per column J != p,
// each forms R[J] = P X[p,J] once for all the tiles of its column. The other
// tiles of all jobs, concatenated (job j's are [first[j], first[j + 1]), in
// row order), are cut into XCDS equal ranges and range x goes to the
// workgroups b = nbase + XCDS i + x, which the dispatcher places on XCD x:
// each XCD reads a job's pivot-row and pivot-column tiles into its own L2
// only for the jobs (or rows of a job) it holds. A tile workgroup that needs
// R[J] waits for its column workgroup's flag; the dispatcher starts the
// workgroups of each XCD in index order, so every column workgroup has
// started before any tile workgroup of its XCD, and column workgroups wait
// for nothing.
constexpr int XCDS = 8;
template <typename T> struct FBatch {
  int njobs;
  int crow;                // ceil(njobs / XCDS)
  int total;               // first[njobs]: the tile workgroups of all jobs
  int rtotal;              // rfirst[njobs]: the column workgroups of all jobs
  int nbase;               // first tile workgroup: XCDS * (crow + ceil(rtotal / XCDS))
  unsigned cap;            // polls of a bounded wait before it gives up
  int first[MAXJ + 1];
  int rfirst[MAXJ + 1];
  FJob<T> job[MAXJ];
};
// Tiles per launch of a job, and its critical tile: (p+1, p+1), or (0, 0)
// at the last step of a block (and the only tile of the initial pivot).
//
// Symmetric steps. D A is complex symmetric for the diagonal stretch matrix D
// (SC-PML), so D_k S_k is symmetric for every Schur complement S_k, and the
// working matrix M of the Gauss-Jordan sweep keeps
//   M[j,i] = s_i s_j (D_i / D_j) M[i,j],  s = -1 on pivots already taken, +1 else
// (D restricted to one block: the ratio depends on y only). A job with dsym
// set computes only the tiles I <= J of each step, reads X[p,J] (J < p) and
// X[I,p] (I > p) through that relation, and at the last step of a block also
// writes the tiles below the diagonal, so G_k is stored whole.
template <typename T>
__host__ __device__ inline int step_tiles(const FJob<T> &J) {
  return J.k < 0 ? 1 : (J.dsym ? J.nt * (J.nt + 1) / 2 : J.nt * J.nt);
}
template <typename T>
__host__ __device__ inline int critical_tile(const FJob<T> &J) {
  return (J.k >= 0 && J.p + 1 < J.nt) ? (J.p + 1) * (J.nt + 1) : 0;
}
// Column workgroups: a symmetric job (after its initial pivot, with a
// column buffer) forms R[J] = P X[p,J] for every J != p in workgroups of
// their own; its tiles of row p (J > p) are then those workgroups' output.
template <typename T>
__host__ __device__ inline bool uses_columns(const FJob<T> &J) {
  return J.dsym && J.rflag && J.k >= 0 && J.nt > 1;
}
template <typename T>
__host__ __device__ inline int column_wgs(const FJob<T> &J) {
  return uses_columns(J) ? J.nt - 1 : 0;
}
// Tile workgroups of a job: its tiles but the critical one and, with column
// workgroups, those of row p right of the diagonal.
template <typename T>
__host__ __device__ inline int tile_wgs(const FJob<T> &J) {
  return step_tiles(J) - 1 - (uses_columns(J) ? J.nt - 1 - J.p : 0);
}
// Tile (I * nt + J) of the u-th tile I <= J of an nt x nt tile grid in row
// order. Row I starts at S(I) = I nt - I (I - 1) / 2: I is the larger root
// of S(I) = u, rounded down, then corrected by one either way for rounding.
__device__ inline int upper_tile(int nt, int u) {
  const float h = nt + 0.5f;
  int I = (int)(h - sqrtf(h * h - 2.0f * u));
  if (I * nt - I * (I - 1) / 2 > u) --I;
  else if (I + 1 < nt && (I + 1) * nt - (I + 1) * I / 2 <= u) ++I;
  return I * nt + u - (I * nt - I * (I - 1) / 2) + I;
}
// Tile of the r-th tile workgroup of a job.
template <typename T>
__device__ inline int step_tile_of(const FJob<T> &J, int r) {
  const int c = critical_tile(J), nt = J.nt, p = J.p;
  if (!J.dsym) return r < c ? r : r + 1;
  if (uses_columns(J)) {
    // row order with (p, p+1) .. (p, nt-1) and the critical (p+1, p+1) taken
    // out (one run after (p, p) at S(p)); at p = nt - 1 only (0, 0) is out
    if (p == nt - 1) return upper_tile(nt, r + 1);
    const int sp = p * nt - p * (p - 1) / 2;
    return upper_tile(nt, r <= sp ? r : r + nt - p);
  }
  // row order with the critical tile (on the diagonal) taken out
  const int ci = c / nt, cu = ci * nt - ci * (ci - 1) / 2;
  return upper_tile(nt, r < cu ? r : r + 1);
}
template <typename T>
inline int batch_layout(FBatch<T> &b) {
  int w = 0, r = 0;
  for (int j = 0; j < b.njobs; ++j) {
    b.first[j] = w;
    b.rfirst[j] = r;
    w += tile_wgs(b.job[j]);
    r += column_wgs(b.job[j]);
  }
  b.first[b.njobs] = w;
  b.rfirst[b.njobs] = r;
  b.total = w;
  b.rtotal = r;
  b.crow = (b.njobs + XCDS - 1) / XCDS;
  b.nbase = XCDS * (b.crow + (r + XCDS - 1) / XCDS);
  // the column waits keep the default bound: the spin-cap setting shortens
  // the persistent kernels' waits, which have a fallback; these have none
  b.cap = SPIN_CAP;
  return b.nbase + XCDS * ((w + XCDS - 1) / XCDS);
}

// An empty asm statement that takes a job's fields in scalar registers: the
// reads of all of them issue before it and complete in one wait, instead of
// one round trip each where the compiler would wait on them behind the
// branches that depend on earlier ones.
template <typename T>
__device__ inline void in_sgprs(const FJob<T> &J) {
  asm volatile("" ::"s"(J.op), "s"(J.G), "s"(J.W), "s"(J.Pinv), "s"(J.status), "s"(J.dense), "s"(J.dsym),
               "s"(J.rbuf), "s"(J.rflag), "s"(J.m), "s"(J.n), "s"(J.np), "s"(J.nt), "s"(J.k), "s"(J.p),
               "s"(J.dir), "s"(J.c));
}

// Four consecutive elements of a 16-byte aligned row from global memory.
__device__ inline void ld_row4(const float *p, float (&v)[4]) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 x = *reinterpret_cast<const f4 *>(p);
  v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
}
__device__ inline void ld_row4(const double *p, double (&v)[4]) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 x = reinterpret_cast<const d2 *>(p)[0], y = reinterpret_cast<const d2 *>(p)[1];
  v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
}

// Four consecutive elements to global memory as 16-byte stores. Plain
// (write-back) stores: every reader of these is a later launch, and with
// non-temporal stores saturated step launches took about 10% longer.
__device__ inline void store4(float *p, const float *v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  *reinterpret_cast<f4 *>(p) = f4{v[0], v[1], v[2], v[3]};
}
__device__ inline void store4(double *p, const double *v) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  reinterpret_cast<d2 *>(p)[0] = d2{v[0], v[1]};
  reinterpret_cast<d2 *>(p)[1] = d2{v[2], v[3]};
}

// Element (a, b) of the working matrix of block k before step p. For p == 0
// this is the Schur complement of block k (identity on the padding):
// S_k = T_k - XM_k G_{k-1} XP_{k-1} on the top chain,
// S'_k = T_k - XP_k G_{k+1} XM_{k+1} on the bottom chain, both terms at the
// twist block.
template <typename T>
__device__ inline void x_elem(const FJob<T> &J, int k, int p, const T *src, int a, int b,
                              T &re, T &im) {
  const size_t np2 = (size_t)J.np * J.np;
  if (p > 0) {
    re = src[(size_t)a * J.np + b];
    im = src[np2 + (size_t)a * J.np + b];
    return;
  }
  const int n = J.n;
  if (a >= n || b >= n) {
    re = (a == b) ? T(1) : T(0);
    im = T(0);
    return;
  }
  if (J.dense) {
    re = J.dense[(size_t)a * J.np + b];
    im = J.dense[np2 + (size_t)a * J.np + b];
    return;
  }
  const size_t M = (size_t)J.m * n;
  const size_t pa = (size_t)k * n + a;
  double vr = 0, vi = 0;
  if (a == b) { cd c = ld_c(J.op, FDFD_OP_C * M + pa); vr += c.r; vi += c.i; }
  if (b == (a + 1) % n) { cd c = ld_c(J.op, FDFD_OP_YP * M + pa); vr += c.r; vi += c.i; }
  if (b == (a + n - 1) % n) { cd c = ld_c(J.op, FDFD_OP_YM * M + pa); vr += c.r; vi += c.i; }
  // the eliminated neighbours: the block above (top chain), the block
  // below (bottom chain, and the twist block of the top chain)
  if (J.dir > 0 && k > 0) {
    const T *Gp = J.G + (size_t)(k - 1) * 2 * np2;
    cd gv = {(double)Gp[(size_t)a * J.np + b], (double)Gp[np2 + (size_t)a * J.np + b]};
    cd l = ld_c(J.op, FDFD_OP_XM * M + pa);
    cd u = ld_c(J.op, FDFD_OP_XP * M + (size_t)(k - 1) * n + b);
    cd s = cmul(cmul(l, gv), u);
    vr -= s.r;
    vi -= s.i;
  }
  if (J.dir < 0 ? k + 1 < J.m : (k == J.c && J.c < J.m - 1)) {
    const T *Gn = J.G + (size_t)(k + 1) * 2 * np2;
    cd gv = {(double)Gn[(size_t)a * J.np + b], (double)Gn[np2 + (size_t)a * J.np + b]};
    cd u = ld_c(J.op, FDFD_OP_XP * M + pa);
    cd l = ld_c(J.op, FDFD_OP_XM * M + (size_t)(k + 1) * n + b);
    cd s = cmul(cmul(u, gv), l);
    vr -= s.r;
    vi -= s.i;
  }
  re = (T)vr;
  im = (T)vi;
}

// Exchange buffers of the tile inverses: rows padded to a 16-byte multiple so
// that a thread's four consecutive elements move as one 16-byte LDS access
// (complex64: 36 floats; FP64: 34 doubles). Two of them overlay the start of
// the workgroup's LDS block while the tile itself is held in registers.
template <typename T> constexpr int inv_stride() { return sizeof(T) == 8 ? 34 : 36; }
template <typename T> struct InvX {
  T r[TB][inv_stride<T>()];
  T i[TB][inv_stride<T>()];
};

// LDS of the persistent factorization: three complex tiles. R = P X[p,J]
// replaces X[p,J] once every wave has read it.
template <typename T> struct Lds {
  T Pr[TB][TBP], Pi[TB][TBP];  // pivot inverse
  T Ar[TB][TBP], Ai[TB][TBP];  // X[p,J], then R = P X[p,J]; tile to invert / inverse
  T Br[TB][TBP], Bi[TB][TBP];  // X[I,p]
  T mag[2][TB];                // pivot candidates (double-buffered)
  int perm[TB];
  int colmap[TB];
};

// LDS of the step kernel: two complex tiles (18.5 KB for complex64, eight
// workgroups per CU). P shares the B buffer with X[I,p], which waits in
// registers until R = P X[p,J] is formed; the pad makes room for the tile
// inverse's two exchange buffers.
template <typename T> constexpr int step_lds_pad() {
  return 2 * (int)sizeof(InvX<T>) > 4 * TB * TBP * (int)sizeof(T)
             ? 2 * (int)sizeof(InvX<T>) - 4 * TB * TBP * (int)sizeof(T) : 16;
}
template <typename T> struct LdsStep {
  T Ar[TB][TBP], Ai[TB][TBP];  // X[p,J], then R; or P (column p); tile to invert / inverse
  T Br[TB][TBP], Bi[TB][TBP];  // P, then X[I,p]
  unsigned char pad[step_lds_pad<T>()];
  alignas(16) T dmir[4][TB];   // D of the tile's rows (re, im), 1/D of its columns (re, im)
  T mag[2][TB];
  int perm[TB];
  int colmap[TB];
};

template <typename T>
__device__ inline void load_tile(const FJob<T> &J, int k, int p, const T *src, int ti, int tj,
                                 T (*dr)[TBP], T (*di)[TBP]) {
  for (int e = threadIdx.x; e < TB * TB; e += NT) {
    const int a = e / TB, b = e % TB;
    x_elem(J, k, p, src, ti * TB + a, tj * TB + b, dr[a][b], di[a][b]);
  }
}

// c += sign * A[quadrant rows, :] * B[:, quadrant cols] (complex, on MFMA).
// Lane group g = lane / 16 supplies index k = 2 s + g / 2 + 16 (g % 2) of
// the inner dimension at MFMA s (s = 0..7; every k once): within each half
// wave the rows r + 16 (g % 2) of A and the columns of B then fall in 32
// distinct LDS banks under the 33-word row pitch (consecutive k, as k = 4 s
// + g, put rows r and r + 1 of the two groups on one bank).
template <typename T>
__device__ inline void cgemm_quad(typename Mfma<T>::v4 &cr, typename Mfma<T>::v4 &ci,
                                  const T (*Ar)[TBP], const T (*Ai)[TBP], const T (*Br)[TBP],
                                  const T (*Bi)[TBP], int qi, int qj, int lane, T sign) {
  static_assert(TB == 32 && TBP == 33, "the inner-index order assumes 32-wide tiles, 33-word rows");
  const int r = lane & 15, kq = lane >> 4;
  const int k0 = (kq >> 1) + 16 * (kq & 1);
#pragma unroll
  for (int s = 0; s < TB / 4; ++s) {
    const int k = 2 * s + k0;
    const T ar = sign * Ar[qi * 16 + r][k], ai = sign * Ai[qi * 16 + r][k];
    const T br = Br[k][qj * 16 + r], bi = Bi[k][qj * 16 + r];
    cr = Mfma<T>::mac(ar, br, cr);
    cr = Mfma<T>::mac(-ai, bi, cr);
    ci = Mfma<T>::mac(ar, bi, ci);
    ci = Mfma<T>::mac(ai, br, ci);
  }
}

// 1/(zr + i zi) from hardware reciprocals with one Newton step each (no IEEE
// division sequences on the critical path); scaled by max(|zr|, |zi|) so
// the squared modulus cannot overflow.
__device__ inline float rcp_nr(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return r * (2.0f - x * r);
}
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * (2.0 - x * r);
  return r * (2.0 - x * r);
}
template <typename T>
__device__ inline void fast_crecip(T zr, T zi, T &outr, T &outi) {
  const T sc = fmax(fabs(zr), fabs(zi));
  const T rs = rcp_nr(sc);
  const T ar = zr * rs, ai = zi * rs;
  const T rd = rcp_nr(ar * ar + ai * ai) * rs;
  outr = ar * rd;
  outi = -ai * rd;
}

__device__ inline void lds_st4(float *p, const float (&v)[4]) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  *reinterpret_cast<f4 *>(p) = f4{v[0], v[1], v[2], v[3]};
}
__device__ inline void lds_st4(double *p, const double (&v)[4]) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  reinterpret_cast<d2 *>(p)[0] = d2{v[0], v[1]};
  reinterpret_cast<d2 *>(p)[1] = d2{v[2], v[3]};
}
__device__ inline void lds_ld4(const float *p, float (&v)[4]) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 x = *reinterpret_cast<const f4 *>(p);
  v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
}
__device__ inline void lds_ld4(const double *p, double (&v)[4]) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 x = reinterpret_cast<const d2 *>(p)[0], y = reinterpret_cast<const d2 *>(p)[1];
  v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
}
template <typename T>
__device__ inline void lds_ld2(const T *p, T &a, T &b) {
  typedef T t2 __attribute__((ext_vector_type(2)));
  const t2 x = *reinterpret_cast<const t2 *>(p);
  a = x.x;
  b = x.y;
}
// 2^-e with 2^(e-1) <= |x| < 2^e: an exact scale that brings x to [0.5, 1)
__device__ inline float exp_scale(float x) { return ldexpf(1.0f, -__builtin_amdgcn_frexp_expf(x)); }
__device__ inline double exp_scale(double x) { return ldexp(1.0, -__builtin_amdgcn_frexp_exp(x)); }

// Inverse of the TB x TB tile in (Ar, Ai), in place, by Gauss-Jordan with
// 2 x 2 diagonal block pivots (the default; the tile is a diagonal block of
// a Schur complement of the PML-damped operator): TB / 2 steps, each inverting the pivot block
// in closed form and applying a rank-2 update. Thread t keeps row t/8,
// columns 4(t%8)..+3 in registers and publishes them each step to one of two
// LDS exchange buffers, so one barrier per step orders the exchange; a step
// reads the pivot block, its row's pivot columns and the pivot rows at its
// columns with ten LDS accesses, and the pivot rows are handled by the same
// update formula as the other rows (no divergent branch). Scales are exact
// powers of two.
template <typename T, typename LT>
__device__ void tile_inverse_diag(LT &L, int *status) {
  static_assert(TB * TB == 4 * NT, "4 elements per thread");
  static_assert(2 * sizeof(InvX<T>) <= offsetof(LT, mag), "exchange buffers overlay the tiles only");
  const int tid = threadIdx.x;
  const int i = tid >> 3, c0 = (tid & 7) * 4;
  T ar[4], ai[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) { ar[q] = L.Ar[i][c0 + q]; ai[q] = L.Ai[i][c0 + q]; }
  __syncthreads();   // every row is in registers before the exchange buffers cover Ar, Ai
  InvX<T> *X = reinterpret_cast<InvX<T> *>(&L);
  bool bad = false;
#pragma unroll 2
  for (int j = 0; j < TB; j += 2) {
    InvX<T> &B = X[(j >> 1) & 1];
    lds_st4(&B.r[i][c0], ar);
    lds_st4(&B.i[i][c0], ai);
    __syncthreads();
    // pivot block, this row's pivot columns, the pivot rows at these columns
    T p00r, p01r, p00i, p01i, p10r, p11r, p10i, p11i, f0r, f1r, f0i, f1i;
    lds_ld2(&B.r[j][j], p00r, p01r);
    lds_ld2(&B.i[j][j], p00i, p01i);
    lds_ld2(&B.r[j + 1][j], p10r, p11r);
    lds_ld2(&B.i[j + 1][j], p10i, p11i);
    lds_ld2(&B.r[i][j], f0r, f1r);
    lds_ld2(&B.i[i][j], f0i, f1i);
    T h0r[4], h0i[4], h1r[4], h1i[4];
    lds_ld4(&B.r[j][c0], h0r);
    lds_ld4(&B.i[j][c0], h0i);
    lds_ld4(&B.r[j + 1][c0], h1r);
    lds_ld4(&B.i[j + 1][c0], h1i);
    // Q = P^-1 = adj(P') sc / det(P') with P' = sc P, sc an exact power of two
    // near 1 / max |component| (det P itself would overflow float for Hz,
    // whose entries reach ~1e20)
    const T mx = fmax(fmax(fmax(fabs(p00r), fabs(p00i)), fmax(fabs(p01r), fabs(p01i))),
                      fmax(fmax(fabs(p10r), fabs(p10i)), fmax(fabs(p11r), fabs(p11i))));
    const T sc = exp_scale(mx);
    const T a00r = p00r * sc, a00i = p00i * sc, a01r = p01r * sc, a01i = p01i * sc;
    const T a10r = p10r * sc, a10i = p10i * sc, a11r = p11r * sc, a11i = p11i * sc;
    const T dr = (a00r * a11r - a00i * a11i) - (a01r * a10r - a01i * a10i);
    const T di = (a00r * a11i + a00i * a11r) - (a01r * a10i + a01i * a10r);
    if (!(fabs(dr) + fabs(di) > T(0)) || !isfinite(dr) || !isfinite(di) || !(mx > T(0)) || !isfinite(mx))
      bad = true;
    // 1 / det: (dr - i di) / |det|^2 with det brought to [0.5, 1) by an exact scale
    const T ds = exp_scale(fmax(fabs(dr), fabs(di)));
    const T br = dr * ds, bi = di * ds;
    const T rm = rcp_nr(br * br + bi * bi) * ds * sc;
    const T er = br * rm, ei = -bi * rm;
    const T q00r = a11r * er - a11i * ei, q00i = a11r * ei + a11i * er;
    const T q01r = -(a01r * er - a01i * ei), q01i = -(a01r * ei + a01i * er);
    const T q10r = -(a10r * er - a10i * ei), q10i = -(a10r * ei + a10i * er);
    const T q11r = a00r * er - a00i * ei, q11i = a00r * ei + a00i * er;
    // new row = base - g0 h0 - g1 h1: off the pivot rows base = row and
    // g = [f0 f1] Q; on them base = 0 and g = -(row of Q), so the pivot rows
    // become Q [h0; h1] without cancellation
    T g0r = f0r * q00r - f0i * q00i + f1r * q10r - f1i * q10i;
    T g0i = f0r * q00i + f0i * q00r + f1r * q10i + f1i * q10r;
    T g1r = f0r * q01r - f0i * q01i + f1r * q11r - f1i * q11i;
    T g1i = f0r * q01i + f0i * q01r + f1r * q11i + f1i * q11r;
    const bool r0 = i == j, r1 = i == j + 1, prow = r0 || r1;
    g0r = r0 ? -q00r : (r1 ? -q10r : g0r);
    g0i = r0 ? -q00i : (r1 ? -q10i : g0i);
    g1r = r0 ? -q01r : (r1 ? -q11r : g1r);
    g1i = r0 ? -q01i : (r1 ? -q11i : g1i);
    // the pivot columns: Q on the pivot rows, -g elsewhere
    const T cjr = r0 ? q00r : (r1 ? q10r : -g0r), cji = r0 ? q00i : (r1 ? q10i : -g0i);
    const T ckr = r0 ? q01r : (r1 ? q11r : -g1r), cki = r0 ? q01i : (r1 ? q11i : -g1i);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const T xr0 = prow ? T(0) : ar[q], xi0 = prow ? T(0) : ai[q];
      const T xr = xr0 - (g0r * h0r[q] - g0i * h0i[q]) - (g1r * h1r[q] - g1i * h1i[q]);
      const T xi = xi0 - (g0r * h0i[q] + g0i * h0r[q]) - (g1r * h1i[q] + g1i * h1r[q]);
      const int c = c0 + q;
      ar[q] = c == j ? cjr : (c == j + 1 ? ckr : xr);
      ai[q] = c == j ? cji : (c == j + 1 ? cki : xi);
    }
  }
  __syncthreads();   // the last step's reads are done before Ar, Ai are written
#pragma unroll
  for (int q = 0; q < 4; ++q) { L.Ar[i][c0 + q] = ar[q]; L.Ai[i][c0 + q] = ai[q]; }
  if (bad) *status = 1;
  __syncthreads();
}

// Inverse of the TB x TB tile in (Ar, Ai), in place, by Gauss-Jordan with
// partial pivoting inside the tile (the fallback factorization). Thread t
// keeps row t/8, columns 4(t%8)..+3 in registers; each pivot step publishes
// the tile to one of the two exchange buffers (alternating), so one barrier
// per pivot orders the exchange. The row interchanges become one permuted
// write-back.
template <typename T, typename LT>
__device__ void tile_inverse_pivoted(LT &L, int *status) {
  static_assert(TB * TB == 4 * NT, "4 elements per thread");
  static_assert(2 * sizeof(InvX<T>) <= offsetof(LT, mag), "exchange buffers overlay the tiles only");
  const int tid = threadIdx.x, lane = tid & 63;
  const int i = tid >> 3, c0 = (tid & 7) * 4;
  T ar[4], ai[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) { ar[q] = L.Ar[i][c0 + q]; ai[q] = L.Ai[i][c0 + q]; }
  __syncthreads();   // every row is in registers before the exchange buffers cover Ar, Ai
  InvX<T> *X = reinterpret_cast<InvX<T> *>(&L);
  bool bad = false;
  for (int j = 0; j < TB; ++j) {
    InvX<T> &B = X[j & 1];
    lds_st4(&B.r[i][c0], ar);
    lds_st4(&B.i[i][c0], ai);
    if (j >= c0 && j < c0 + 4) {
      const int q = j - c0;
      L.mag[j & 1][i] = (i >= j) ? fabs(ar[q]) + fabs(ai[q]) : T(-1);
    }
    __syncthreads();
    // argmax over the column (ties to the lower row), in every wave
    T m = (lane < TB) ? L.mag[j & 1][lane] : T(-2);
    if (!(m == m)) m = T(-1);  // a NaN candidate is never chosen
    int idx = lane;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const T om = __shfl_xor(m, off);
      const int oi = __shfl_xor(idx, off);
      if (om > m || (om == m && oi < idx)) { m = om; idx = oi; }
    }
    const int piv = (idx >= j && idx < TB) ? idx : j;
    if (tid == 0) L.perm[j] = piv;
    const int er = (i == j) ? piv : (i == piv ? j : i);  // row content after the swap
    if (er != i) {
#pragma unroll
      for (int q = 0; q < 4; ++q) { ar[q] = B.r[er][c0 + q]; ai[q] = B.i[er][c0 + q]; }
    }
    const T pr = B.r[piv][j], pim = B.i[piv][j];
    if (!(fabs(pr) + fabs(pim) > T(0)) || !isfinite(pr) || !isfinite(pim)) bad = true;
    T dr, di;
    crecip<T>(pr, pim, dr, di);
    if (i == j) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (c0 + q == j) { ar[q] = dr; ai[q] = di; }
        else {
          const T xr = ar[q], xi = ai[q];
          ar[q] = xr * dr - xi * di;
          ai[q] = xr * di + xi * dr;
        }
      }
    } else {
      const T fr = B.r[er][j], fi = B.i[er][j];
      const T gr = fr * dr - fi * di, gi = fr * di + fi * dr;  // f * d
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (c0 + q == j) { ar[q] = -gr; ai[q] = -gi; }
        else {
          const T hr = B.r[piv][c0 + q], hi = B.i[piv][c0 + q];
          ar[q] -= gr * hr - gi * hi;
          ai[q] -= gr * hi + gi * hr;
        }
      }
    }
  }
  // TB is even: buffer 0 was last read at step TB - 2, before step TB - 1's barrier
  lds_st4(&X[0].r[i][c0], ar);
  lds_st4(&X[0].i[i][c0], ai);
  if (tid == 0) {
    for (int c = 0; c < TB; ++c) L.colmap[c] = c;
    for (int j = TB - 1; j >= 0; --j) {
      const int pj = L.perm[j], t = L.colmap[j];
      L.colmap[j] = L.colmap[pj];
      L.colmap[pj] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int src = L.colmap[c0 + q];
    ar[q] = X[0].r[i][src];
    ai[q] = X[0].i[i][src];
  }
  __syncthreads();   // every read of buffer 0 is done before Ar, Ai (which it overlays) are written
#pragma unroll
  for (int q = 0; q < 4; ++q) { L.Ar[i][c0 + q] = ar[q]; L.Ai[i][c0 + q] = ai[q]; }
  if (bad) *status = 1;
  __syncthreads();
}

template <typename T, bool PIV, typename LT>
__device__ inline void tile_inverse(LT &L, int *status) {
  if (PIV) tile_inverse_pivoted<T>(L, status);
  else tile_inverse_diag<T>(L, status);
}

template <typename T, typename LT>
__device__ inline void store_pinv(const LT &L, T *slot) {
  // whole rows, 16-byte stores (read by the next launch)
  const int i = threadIdx.x >> 3, c0 = (threadIdx.x & 7) * 4;
  store4(slot + i * TB + c0, &L.Ar[i][c0]);
  store4(slot + TB * TB + i * TB + c0, &L.Ai[i][c0]);
}

// The [0,0] tile of the chain's next block kn = k + dir from G_k[0,0] held
// in (Ar, Ai): T_kn - XM_kn G_k XP_k (top chain) or T_kn - XP_kn G_k XM_k
// (bottom chain); at the twist block also - XP_kn G_{kn+1} XM_{kn+1}, from the
// bottom chain's last block, finished in an earlier launch.
template <typename T, typename LT>
__device__ inline void schur_next_00(const FJob<T> &J, int k, LT &L) {
  const int n = J.n, kn = k + J.dir;
  const size_t M = (size_t)J.m * n, np2 = (size_t)J.np * J.np;
  const bool twist = J.dir > 0 && kn == J.c && J.c < J.m - 1;
  const T *Gb = J.G + (size_t)(kn + 1) * 2 * np2;   // read only at the twist block
  for (int e = threadIdx.x; e < TB * TB; e += NT) {
    const int a = e / TB, b = e % TB;
    double vr, vi;
    if (a >= n || b >= n) { vr = (a == b) ? 1.0 : 0.0; vi = 0.0; }
    else {
      const size_t pa = (size_t)kn * n + a;
      vr = 0; vi = 0;
      if (a == b) { cd c = ld_c(J.op, FDFD_OP_C * M + pa); vr += c.r; vi += c.i; }
      if (b == (a + 1) % n) { cd c = ld_c(J.op, FDFD_OP_YP * M + pa); vr += c.r; vi += c.i; }
      if (b == (a + n - 1) % n) { cd c = ld_c(J.op, FDFD_OP_YM * M + pa); vr += c.r; vi += c.i; }
      cd gv = {(double)L.Ar[a][b], (double)L.Ai[a][b]};
      const cd s = J.dir > 0
          ? cmul(cmul(ld_c(J.op, FDFD_OP_XM * M + pa), gv), ld_c(J.op, FDFD_OP_XP * M + (size_t)k * n + b))
          : cmul(cmul(ld_c(J.op, FDFD_OP_XP * M + pa), gv), ld_c(J.op, FDFD_OP_XM * M + (size_t)k * n + b));
      vr -= s.r; vi -= s.i;
      if (twist) {
        cd g2 = {(double)Gb[(size_t)a * J.np + b], (double)Gb[np2 + (size_t)a * J.np + b]};
        const cd s2 = cmul(cmul(ld_c(J.op, FDFD_OP_XP * M + pa), g2),
                           ld_c(J.op, FDFD_OP_XM * M + (size_t)(kn + 1) * n + b));
        vr -= s2.r; vi -= s2.i;
      }
    }
    L.Ar[a][b] = (T)vr;  // each (a,b) read and written by the same thread
    L.Ai[a][b] = (T)vi;
  }
  __syncthreads();
}

// D over one block from the couplings of block 0: D_{y+1} / D_y =
// A[y, y+1] / A[y+1, y] (the same ratio in every block); 1 on the padding.
template <typename T>
__device__ void block_symmetrizer(const double *op, int m, int n, int np, T *dsym) {
  // D_0 = 1, D_y = D_{y-1} A[y-1,y] / A[y,y-1] for y < n, 1 beyond: a
  // product scan by the first wave, 64 entries per pass (lane l takes the
  // ratio into y = y0 + l; a shuffle scan forms the products within the
  // pass, the last lane's product carries to the next)
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const size_t M = (size_t)m * n;
  cd carry = {1.0, 0.0};
  for (int y0 = 0; y0 < np; y0 += 64) {
    const int y = y0 + lane;
    cd r = {1.0, 0.0};
    if (y >= 1 && y < n) r = cmul(ld_c(op, FDFD_OP_YP * M + y - 1), cinv(ld_c(op, FDFD_OP_YM * M + y)));
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const cd o = {__shfl_up(r.r, off), __shfl_up(r.i, off)};
      if (lane >= off) r = cmul(o, r);
    }
    const cd pd = cmul(carry, r);
    carry = {__shfl(pd.r, 63), __shfl(pd.i, 63)};
    const cd d = y < n ? pd : cd{1.0, 0.0};
    if (y < np) {
      const cd di = cinv(d);
      dsym[y] = (T)d.r;
      dsym[np + y] = (T)d.i;
      dsym[2 * np + y] = (T)di.r;
      dsym[3 * np + y] = (T)di.i;
    }
  }
}
template <typename T>
__device__ void block_symmetrizer(const FJob<T> &J) {
  block_symmetrizer(J.op, J.m, J.n, J.np, J.dsym);
}
// the same D for the corner chains' products, on their own stream
template <typename T>
__global__ void k_block_symmetrizer(const double *op, int m, int n, int np, T *dsym) {
  block_symmetrizer(op, m, n, np, dsym);
}

// Thread 0 waits until column J's flag holds the step's tag (or the job's
// status is set: the factorization has failed anyway); then the whole
// group. After `cap` polls the wait gives up and sets the status to 2, so a
// grid always drains.
template <typename T>
__device__ inline void wait_column(const FJob<T> &J, int col, unsigned tag, unsigned cap) {
  if (threadIdx.x == 0) {
    unsigned spins = 0;
    while (ld_xcd(J.rflag + col) < tag) {
      if ((++spins & 63) == 0 && ld_xcd(J.status)) break;
      if (spins > cap) {
        st_xcd(J.status, 2);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// The same wait by every wave on its own (all lanes read the one word), so
// a wave goes on to read R[J] without a barrier that would also wait for
// the group's other reads in flight.
template <typename T>
__device__ inline void wave_wait_column(const FJob<T> &J, int col, unsigned tag, unsigned cap) {
  unsigned spins = 0;
  while (ld_xcd(J.rflag + col) < tag) {
    if ((++spins & 63) == 0 && ld_xcd(J.status)) break;
    if (spins > cap) {
      if ((threadIdx.x & 63) == 0) st_xcd(J.status, 2);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Row a, columns b0..b0+3 of R[col] once its flag is up, as plain 16-byte
// loads. They may be served by this XCD's L2: a line of R[col] reaches an
// L2 only through a read made after the flag was seen raised, when the
// column workgroup's write-through stores have completed (caches hold no
// line of the buffer from an earlier launch), so the first reader on an XCD
// fetches it from memory and the others hit. The empty asm keeps the
// compiler from moving the loads above the flag reads; it adds no wait on
// the loads already in flight.
template <typename T>
__device__ inline void ld_column(const FJob<T> &J, int col, int a, int b0, T (&vr)[TB * TB / NT],
                                 T (&vi)[TB * TB / NT]) {
  asm volatile("" ::: "memory");
  const T *R = J.rbuf + (size_t)col * 2 * TB * TB;
  ld_row4(R + a * TB + b0, vr);
  ld_row4(R + TB * TB + a * TB + b0, vi);
}

// The column workgroup r of a job: R[J] = P X[p,J] for column J = r (r < p)
// or r + 1, into the job's column buffer (agent-coherent stores, then the
// flag); right of the diagonal R[J] is also the step's result for tile
// (p, J). X[p,J] is read like the tiles' X[p,J]: as the previous output,
// through the symmetry for J < p, or formed from the operator at p = 0.
template <typename T>
__device__ void column_step(const FJob<T> &J, LdsStep<T> &L, int r, const T *src, T *dst, const T *Pslot,
                            unsigned tag) {
  const int p = J.p, col = r < p ? r : r + 1;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, qi = w >> 1, qj = w & 1;
  const size_t np2 = (size_t)J.np * J.np;
  constexpr int PER = TB * TB / NT;
  const int a = tid >> 3, b0 = (tid & 7) * 4;
  if (p > 0) {
    const bool mir = col < p;   // X[p,J] = -(D[32 J + a] / D[32 p + b]) X[J,p]^T
    T pr[PER], pim[PER], xr[PER], xi[PER], er[PER], ei[PER], dr = 0, di = 0;
    ld_row4(Pslot + a * TB + b0, pr);
    ld_row4(Pslot + TB * TB + a * TB + b0, pim);
    const size_t ia = mir ? (size_t)(col * TB + a) * J.np + p * TB + b0 : (size_t)(p * TB + a) * J.np + col * TB + b0;
    ld_row4(src + ia, xr);
    ld_row4(src + np2 + ia, xi);
    if (mir) {
      dr = J.dsym[col * TB + a];
      di = J.dsym[J.np + col * TB + a];
      ld_row4(J.dsym + 2 * J.np + p * TB + b0, er);
      ld_row4(J.dsym + 3 * J.np + p * TB + b0, ei);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      L.Br[a][b0 + u] = pr[u];
      L.Bi[a][b0 + u] = pim[u];
      if (mir) {
        const T rr = -(dr * er[u] - di * ei[u]), ri = -(dr * ei[u] + di * er[u]);
        L.Ar[b0 + u][a] = rr * xr[u] - ri * xi[u];
        L.Ai[b0 + u][a] = rr * xi[u] + ri * xr[u];
      } else {
        L.Ar[a][b0 + u] = xr[u];
        L.Ai[a][b0 + u] = xi[u];
      }
    }
  } else {
    for (int e = tid; e < TB * TB; e += NT) {
      L.Br[e / TB][e % TB] = Pslot[e];
      L.Bi[e / TB][e % TB] = Pslot[TB * TB + e];
    }
    load_tile(J, J.k, 0, src, 0, col, L.Ar, L.Ai);
  }
  __syncthreads();
  typename Mfma<T>::v4 rr = {0, 0, 0, 0}, ri = {0, 0, 0, 0};
  cgemm_quad<T>(rr, ri, L.Br, L.Bi, L.Ar, L.Ai, qi, qj, lane, T(1));
  T *R = J.rbuf + (size_t)col * 2 * TB * TB;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int ra = qi * 16 + Mfma<T>::crow(lane, t), rb = qj * 16 + (lane & 15);
    st_xcd(R + ra * TB + rb, rr[t]);
    st_xcd(R + TB * TB + ra * TB + rb, ri[t]);
  }
  if (col > p) {
    // tile (p, J) of the step's output: whole rows through LDS
    __syncthreads();   // every wave has read X[p,J]
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int ra = qi * 16 + Mfma<T>::crow(lane, t), rb = qj * 16 + (lane & 15);
      L.Ar[ra][rb] = rr[t];
      L.Ai[ra][rb] = ri[t];
    }
    __syncthreads();
    const size_t idx = (size_t)(p * TB + a) * J.np + col * TB + b0;
    store4(dst + idx, &L.Ar[a][b0]);
    store4(dst + np2 + idx, &L.Ai[a][b0]);
  }
  stores_complete();
  __syncthreads();
  if (tid == 0) st_xcd(J.rflag + col, tag);
}

// One Gauss-Jordan step (k, p) for every job of the batch; one output tile
// (I, J) per workgroup:
//   I == p, J == p : P                     (P = inverse of X[p,p])
//   I == p, J != p : P X[p,J]
//   I != p, J == p : -X[I,p] P
//   otherwise      : X[I,J] - X[I,p] (P X[p,J])
// The workgroup owning tile (p+1, p+1) also inverts its result for the next
// step; at the last step the owner of (0,0) forms and inverts S_{k+1}[0,0].
// The bound asks for at least six waves per SIMD (without it the register
// allocation held the kernel to four). With 18.5 KB of LDS, eight complex64
// workgroups fit a CU when a lane needs at most 64 registers; FP64 (35 KB)
// is held to four by its LDS.
template <typename T, bool PIV>
__global__ void __launch_bounds__(NT, 6) k_gj_step(FBatch<T> batch) {
  __shared__ LdsStep<T> L;
  // The job of this workgroup. The batch lives in the kernel arguments. Lane
  // j of each wave reads first[j] and rfirst[j] (vector reads beside the
  // header's scalar reads); a ballot then gives the last job whose
  // workgroups start at or before this one, and the job's fields follow in
  // one more round trip. (A scan with the job index in a scalar register
  // waited for one read per job; an unrolled scalar search kept the CU's
  // scalar unit busy.)
  const int lane0 = threadIdx.x & 63;
  // a critical workgroup's job is job[blockIdx]: its fields are read with
  // the header (clamped, so every workgroup reads inside the batch)
  const FJob<T> Jc = batch.job[(int)blockIdx.x < MAXJ ? (int)blockIdx.x : MAXJ - 1];
  const int njobs = batch.njobs, crow = batch.crow, wtot = batch.total, rtot = batch.rtotal;
  const int nbase = batch.nbase;
  const int fl = batch.first[lane0 < MAXJ ? lane0 : 0], rl = batch.rfirst[lane0 < MAXJ ? lane0 : 0];
  in_sgprs(Jc);
  asm volatile("" ::"s"(njobs), "s"(crow), "s"(wtot), "s"(rtot), "s"(nbase));
  const int bid = blockIdx.x, xcd = bid % XCDS, row = (bid - nbase) / XCDS;
  const bool crit = bid < njobs;
  const int cb = bid - XCDS * crow;   // column workgroup index
  const bool colwg = cb >= 0 && cb < rtot;
  const int g = (xcd * wtot) / XCDS + row;
  const int key = colwg ? rl : fl, want = colwg ? cb : g;
  const unsigned long long hit = __ballot(lane0 < njobs && key <= want);
  const int jn = hit ? 63 - __clzll(hit) : 0;
  const int base = __builtin_amdgcn_readlane(key, jn);
  // idle slots beside the critical and the column workgroups, and the end
  // of each range
  if (!crit && !colwg && (bid < nbase || g >= ((xcd + 1) * wtot) / XCDS)) return;
  FJob<T> Jv;
  if (crit) Jv = Jc;
  else Jv = batch.job[jn];
  in_sgprs(Jv);
  const FJob<T> &J = Jv;
  // the critical workgroups bound the launch, and the column workgroups
  // gate the tile workgroups of their columns: their waves issue first on
  // the CUs they share with the others
  if (crit || colwg) __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int qi = w >> 1, qj = w & 1;
  const size_t np2 = (size_t)J.np * J.np;
  const int nt = J.nt;

  if (J.k < 0) {  // initial pivot: inverse of T_k0[0,0], k0 the chain's first block
    if (crit) {
      if (J.dsym && J.dir > 0) block_symmetrizer(J);
      load_tile(J, J.dir > 0 ? 0 : J.m - 1, 0, (const T *)nullptr, 0, 0, L.Ar, L.Ai);
      __syncthreads();
      tile_inverse<T, PIV>(L, J.status);
      store_pinv(L, J.Pinv);  // slot of global step 0
    }
    return;
  }
  const int k = J.k, p = J.p;
  T *Gk = J.G + (size_t)k * 2 * np2;
  const T *src = (p == 0) ? nullptr : (((nt - p) % 2 == 0) ? Gk : J.W);
  T *dst = ((nt - 1 - p) % 2 == 0) ? Gk : J.W;
  const long step = (long)(J.dir > 0 ? k : J.m - 1 - k) * nt + p;   // of this chain
  const T *Pslot = J.Pinv + (size_t)(step & 1) * 2 * TB * TB;
  T *Pnext = J.Pinv + (size_t)((step + 1) & 1) * 2 * TB * TB;
  if (colwg) {
    column_step<T>(J, L, cb - base, src, dst, Pslot, (unsigned)(step + 1));
    return;
  }
  const int local = crit ? critical_tile(J) : step_tile_of(J, g - base);
  const int tj = local % nt, ti = local / nt;
  // a tile off row and column p takes R[J] from its column workgroup
  const bool with_r = !crit && uses_columns(J) && tj != p && ti != p;

  // every global read of this tile's update, issued together. P goes to the
  // B buffer while X[p,J] is in the A buffer (R = P X[p,J]), or to the A
  // buffer when the tile has no X[p,J] (column p: -X[I,p] P); X[I,p] takes
  // the B buffer once R has replaced X[p,J].
  typename Mfma<T>::v4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  const bool has_a = tj != p, has_b = ti != p;
  // the reads of P and X[p,J]: not for a tile that takes R[J]
  const bool own_r = !with_r;
  T (*Pbr)[TBP] = has_a ? L.Br : L.Ar;
  T (*Pbi)[TBP] = has_a ? L.Bi : L.Ai;
  constexpr int PER = TB * TB / NT;
  T br[PER], bi[PER];
  // symmetric jobs read X[p,J] (J < p) as tile (J,p), X[I,p] (I > p) as (p,I)
  const bool mirA = own_r && p > 0 && J.dsym && tj < p, mirB = p > 0 && J.dsym && ti > p;
  if (p > 0 && with_r) {
    // straight-line, so the reads issue together: the column's flag first
    // (its answer comes back while the tile reads are in flight), X[I,p]
    // (through the symmetry when I > p, with its symmetrizer entries read
    // either way), the tile itself; then R[J] once the flag is up
    const int a = tid >> 3, b0 = (tid & 7) * 4;
    const unsigned tag = (unsigned)(step + 1);
    const unsigned flag0 = ld_xcd(J.rflag + tj);
    const size_t ib = mirB ? (size_t)(p * TB + a) * J.np + ti * TB + b0
                           : (size_t)(ti * TB + a) * J.np + p * TB + b0;
    ld_row4(src + ib, br);
    ld_row4(src + np2 + ib, bi);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int ra = qi * 16 + Mfma<T>::crow(lane, t), rb = qj * 16 + (lane & 15);
      const size_t idx = (size_t)(ti * TB + ra) * J.np + tj * TB + rb;
      cr[t] = src[idx];
      ci[t] = src[np2 + idx];
    }
    T er[PER], ei[PER];
    const T dr = J.dsym[p * TB + a], di = J.dsym[J.np + p * TB + a];
    ld_row4(J.dsym + 2 * J.np + ti * TB + b0, er);
    ld_row4(J.dsym + 3 * J.np + ti * TB + b0, ei);
    if (flag0 < tag) wave_wait_column(J, tj, tag, batch.cap);
    T rbr[PER], rbi[PER];
    ld_column(J, tj, a, b0, rbr, rbi);
    // X[I,p][b][a] = (D[32 p + a] / D[32 I + b]) X[p,I][a][b] when mirrored
    T *Bflat_r = &L.Br[0][0], *Bflat_i = &L.Bi[0][0];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const T sr = dr * er[u] - di * ei[u], si = dr * ei[u] + di * er[u];
      const T xr = mirB ? sr * br[u] - si * bi[u] : br[u];
      const T xi = mirB ? sr * bi[u] + si * br[u] : bi[u];
      const int at = mirB ? (b0 + u) * TBP + a : a * TBP + b0 + u;
      Bflat_r[at] = xr;
      Bflat_i[at] = xi;
      L.Ar[a][b0 + u] = rbr[u];
      L.Ai[a][b0 + u] = rbi[u];
    }
  } else if (p > 0) {
    // plain loads of the previous launch's output, all of them before any
    // use: thread t reads four consecutive elements of row t/8 of each tile
    // as one 16-byte access per plane
    T pr[PER], pim[PER], ar[PER], ai[PER];
    const int a = tid >> 3, b0 = (tid & 7) * 4;
    if (own_r) {
      ld_row4(Pslot + a * TB + b0, pr);
      ld_row4(Pslot + TB * TB + a * TB + b0, pim);
    }
    const size_t ia = mirA ? (size_t)(tj * TB + a) * J.np + p * TB + b0
                           : (size_t)(p * TB + a) * J.np + tj * TB + b0;
    const size_t ib = mirB ? (size_t)(p * TB + a) * J.np + ti * TB + b0
                           : (size_t)(ti * TB + a) * J.np + p * TB + b0;
    if (has_a && own_r) {
      ld_row4(src + ia, ar);
      ld_row4(src + np2 + ia, ai);
    }
    if (has_b) {
      ld_row4(src + ib, br);
      ld_row4(src + np2 + ib, bi);
    }
    if (has_a && has_b) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int a = qi * 16 + Mfma<T>::crow(lane, t), b = qj * 16 + (lane & 15);
        const size_t idx = (size_t)(ti * TB + a) * J.np + tj * TB + b;
        cr[t] = src[idx];
        ci[t] = src[np2 + idx];
      }
    }
    // the symmetrizer entries of a mirrored tile, read with the tiles: D of
    // row 32 R + a and 1/D of columns 32 C + b0.. (R, C = J, p for X[p,J];
    // p, I for X[I,p]; at most one of the two is mirrored when I <= J)
    T dr = 0, di = 0, er[PER], ei[PER];
    if (mirA || mirB) {
      const int R = mirA ? tj : p, C = mirA ? p : ti;
      dr = J.dsym[R * TB + a];
      di = J.dsym[J.np + R * TB + a];
      ld_row4(J.dsym + 2 * J.np + C * TB + b0, er);
      ld_row4(J.dsym + 3 * J.np + C * TB + b0, ei);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int b = b0 + u;
      if (own_r) {
        Pbr[a][b] = pr[u];
        Pbi[a][b] = pim[u];
      }
      // sign * D[32 R + a] / D[32 C + b]
      const T sg = mirA ? T(-1) : T(1);
      const T rr = sg * (dr * er[u] - di * ei[u]), ri = sg * (dr * ei[u] + di * er[u]);
      if (mirA) {  // X[p,J][b][a] = -(D[32 J + a] / D[32 p + b]) X[J,p][a][b]
        L.Ar[b][a] = rr * ar[u] - ri * ai[u];
        L.Ai[b][a] = rr * ai[u] + ri * ar[u];
      } else if (has_a && own_r) {
        L.Ar[a][b] = ar[u];
        L.Ai[a][b] = ai[u];
      }
      if (mirB) {  // X[I,p][b][a] = (D[32 p + a] / D[32 I + b]) X[p,I][a][b], scaled now
        const T xr = rr * br[u] - ri * bi[u], xi = rr * bi[u] + ri * br[u];
        br[u] = xr;
        bi[u] = xi;
      }
    }
  } else {
    // p == 0: the tiles of S_k = T_k - l_k G_{k-1} u_{k-1}, formed on load
    if (own_r)
      for (int e = tid; e < TB * TB; e += NT) {
        Pbr[e / TB][e % TB] = Pslot[e];
        Pbi[e / TB][e % TB] = Pslot[TB * TB + e];
      }
    if (has_a && own_r) load_tile(J, k, p, src, p, tj, L.Ar, L.Ai);   // X[p, J]
    if (has_a && has_b) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int a = qi * 16 + Mfma<T>::crow(lane, t), b = qj * 16 + (lane & 15);
        T xr, xi;
        x_elem(J, k, p, src, ti * TB + a, tj * TB + b, xr, xi);
        cr[t] = xr;
        ci[t] = xi;
      }
    }
  }
  // the symmetrizer entries of the tile's mirror image, stored at a block's
  // last step, read with the tiles and kept in LDS until the store: D of the
  // tile's rows 32 I.. (threads 0-7) and 1/D of its columns 32 J.. (threads
  // 0, 8, .., 248)
  if (J.dsym && p == nt - 1 && ti != tj) {
    if (tid < 8) {
      T fr[4], fi[4];
      ld_row4(J.dsym + ti * TB + 4 * tid, fr);
      ld_row4(J.dsym + J.np + ti * TB + 4 * tid, fi);
      lds_st4(&L.dmir[0][4 * tid], fr);
      lds_st4(&L.dmir[1][4 * tid], fi);
    }
    if ((tid & 7) == 0) {
      L.dmir[2][tid >> 3] = J.dsym[2 * J.np + tj * TB + (tid >> 3)];
      L.dmir[3][tid >> 3] = J.dsym[3 * J.np + tj * TB + (tid >> 3)];
    }
  }
  // X[I, p] into the B buffer: from the registers (p > 0) or formed on load
  auto put_b = [&]() {
    if (p > 0) {
      const int a = tid >> 3, b0 = (tid & 7) * 4;   // the load mapping above
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int b = b0 + u;
        if (mirB) {  // already scaled: transposed into place
          L.Br[b][a] = br[u];
          L.Bi[b][a] = bi[u];
        } else {
          L.Br[a][b] = br[u];
          L.Bi[a][b] = bi[u];
        }
      }
    } else {
      load_tile(J, k, p, src, ti, p, L.Br, L.Bi);
    }
  };
  if (with_r) {
    // R[J] from its column workgroup in the A buffer and X[I,p] in B: read
    // with the tiles when p > 0, here at p = 0
    if (p == 0) {
      wait_column(J, tj, (unsigned)(step + 1), batch.cap);
      const int a = tid >> 3, b0 = (tid & 7) * 4;
      T rbr[PER], rbi[PER];
      ld_column(J, tj, a, b0, rbr, rbi);
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        L.Ar[a][b0 + u] = rbr[u];
        L.Ai[a][b0 + u] = rbi[u];
      }
      put_b();
    }
    __syncthreads();
    cgemm_quad<T>(cr, ci, L.Br, L.Bi, L.Ar, L.Ai, qi, qj, lane, T(-1));
  } else {
  if (!has_a && has_b) put_b();   // column p: P is in the A buffer
  __syncthreads();
  if (has_a) {  // R = P X[p, J]
    typename Mfma<T>::v4 rr = {0, 0, 0, 0}, ri = {0, 0, 0, 0};
    cgemm_quad<T>(rr, ri, L.Br, L.Bi, L.Ar, L.Ai, qi, qj, lane, T(1));
    __syncthreads();  // every wave has read P and X[p, J]: R and X[I, p] take their places
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int a = qi * 16 + Mfma<T>::crow(lane, t), b = qj * 16 + (lane & 15);
      L.Ar[a][b] = rr[t];
      L.Ai[a][b] = ri[t];
    }
    if (has_b) put_b();
    else { cr = rr; ci = ri; }
    __syncthreads();
  }
  if (!has_a && !has_b) {  // (p, p): P, in the A buffer
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int a = qi * 16 + Mfma<T>::crow(lane, t), b = qj * 16 + (lane & 15);
      cr[t] = L.Ar[a][b];
      ci[t] = L.Ai[a][b];
    }
  } else if (has_b) {  // -X[I,p] P (column p) or X[I,J] - X[I,p] R; the right factor is in the A buffer
    cgemm_quad<T>(cr, ci, L.Br, L.Bi, L.Ar, L.Ai, qi, qj, lane, T(-1));
  }
  }
  // the tile goes out through LDS as whole rows (128-byte lines, 16-byte
  // stores); the next launch reads it, possibly on another XCD
  __syncthreads();  // all reads of Ar/Ai, Pr/Pi, Br/Bi are done
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int a = qi * 16 + Mfma<T>::crow(lane, t), b = qj * 16 + (lane & 15);
    L.Ar[a][b] = cr[t];
    L.Ai[a][b] = ci[t];
  }
  __syncthreads();
  const bool next_pivot = (p + 1 < nt) && ti == p + 1 && tj == p + 1;
  const bool next_block = (p == nt - 1) && ti == 0 && tj == 0 && (J.dir > 0 ? k < J.c : k > J.c + 1);
  // the tile of the next pivot is replaced by its inverse at the next step
  // and read by no workgroup before that: its store is skipped
  if (!next_pivot) {
    const int i = tid >> 3, c0 = (tid & 7) * 4;
    const size_t idx = (size_t)(ti * TB + i) * J.np + tj * TB + c0;
    store4(dst + idx, &L.Ar[i][c0]);
    store4(dst + np2 + idx, &L.Ai[i][c0]);
    if (J.dsym && p == nt - 1 && ti != tj) {
      // G_k is symmetric under D: G_k[J,I][i][c] = (D[32 I + c] / D[32 J + i]) G_k[I,J][c][i]
      T mr[4], mi[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // D[32 I + c0 + q] / D[32 J + i]
        const T fr = L.dmir[0][c0 + q], fi = L.dmir[1][c0 + q], gr = L.dmir[2][i], gi = L.dmir[3][i];
        const T rr = fr * gr - fi * gi, ri = fr * gi + fi * gr;
        const T vr = L.Ar[c0 + q][i], vi = L.Ai[c0 + q][i];
        mr[q] = rr * vr - ri * vi;
        mi[q] = rr * vi + ri * vr;
      }
      const size_t mdx = (size_t)(tj * TB + i) * J.np + ti * TB + c0;
      store4(dst + mdx, mr);
      store4(dst + np2 + mdx, mi);
    }
  }
  if (next_pivot || next_block) {
    // the tile is in Ar/Ai; S_{k+1}[0,0] overwrites it only after every
    // thread has read its row for the store above
    if (next_block) {
      __syncthreads();
      schur_next_00(J, k, L);
    }
    tile_inverse<T, PIV>(L, J.status);
    store_pinv(L, Pnext);
  }
}

// ------------------------------------------------ persistent factorization
// One launch factors one operator. Workgroup (I, J) keeps tile (I, J) of the
// working matrix in its MFMA accumulators for the whole factorization. At
// step p of block k the owners of row p and column p publish their tiles
// before updating them, the owner of (p, p) inverts its tile and publishes
// the inverse, and every other workgroup waits for the three tiles it needs.
// Flags hold the block number + 1, so they never need resetting within a
// launch. The exchange buffers alternate between two sets by block parity:
// a workgroup writing set (k & 1) at block k+2 has (through the flags it
// waited on in block k+1) waited for every reader of block k to finish
// block k, so no grid barrier is needed. At the end of block k every
// workgroup writes its tile of G_k and applies the Schur update of block
// k+1 to its registers.
template <typename T> struct PJob {
  const double *op;
  T *G;            // [m][2][np][np]
  T *xrow;         // [2 parity][nt][nt][2][TB*TB]  published X[p, J]
  T *xcol;         // [2 parity][nt][nt][2][TB*TB]  published X[I, p]
  T *pinv;         // [2 parity][nt][2][TB*TB]      published inverses
  unsigned *sync;  // [0] abort, [1] unused, [2, 2+nt) pinv flags, then row / col flags
  int *status;
  int m, n, np, nt;
  unsigned cap;    // polls before a wait gives up
};

template <typename T>
__device__ inline void publish_tile(T *dst, const typename Mfma<T>::v4 &cr, const typename Mfma<T>::v4 &ci,
                                    int qi, int qj, int lane) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int e = (qi * 16 + Mfma<T>::crow(lane, t)) * TB + qj * 16 + (lane & 15);
    st_xcd(dst + e, cr[t]);
    st_xcd(dst + TB * TB + e, ci[t]);
  }
}
// Published tiles are read with every load of the call issued before the
// first LDS store, so their latencies overlap instead of adding up.
template <typename T, int NTILE>
__device__ inline void fetch_tiles(const T *const (&src)[NTILE], T (*const (&dr)[NTILE])[TBP],
                                   T (*const (&di)[NTILE])[TBP]) {
  constexpr int PER = TB * TB / NT;
  T vr[NTILE][PER], vi[NTILE][PER];
#pragma unroll
  for (int t = 0; t < NTILE; ++t)
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * NT;
      vr[t][u] = ld_xcd(src[t] + e);
      vi[t][u] = ld_xcd(src[t] + TB * TB + e);
    }
#pragma unroll
  for (int t = 0; t < NTILE; ++t)
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * NT;
      dr[t][e / TB][e % TB] = vr[t][u];
      di[t][e / TB][e % TB] = vi[t][u];
    }
}
template <typename T>
__device__ inline void fetch_tile(const T *src, T (*dr)[TBP], T (*di)[TBP]) {
  const T *const s1[1] = {src};
  T (*const d1r[1])[TBP] = {dr};
  T (*const d1i[1])[TBP] = {di};
  fetch_tiles<T, 1>(s1, d1r, d1i);
}
template <typename T>
__device__ inline void acc_to_lds(const typename Mfma<T>::v4 &cr, const typename Mfma<T>::v4 &ci,
                                  T (*dr)[TBP], T (*di)[TBP], int qi, int qj, int lane) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int a = qi * 16 + Mfma<T>::crow(lane, t), b = qj * 16 + (lane & 15);
    dr[a][b] = cr[t];
    di[a][b] = ci[t];
  }
}

template <typename T, bool PIV>
__global__ void __launch_bounds__(NT) k_gj_persist(PJob<T> J) {
  __shared__ Lds<T> L;
  const int nt = J.nt;
  const int ti = blockIdx.x / nt, tj = blockIdx.x % nt;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int qi = w >> 1, qj = w & 1;
  const size_t np2 = (size_t)J.np * J.np, tsz = 2 * TB * TB;
  unsigned *abort = J.sync, *fpinv = J.sync + 2;
  unsigned *frow = fpinv + nt, *fcol = frow + nt * nt;
  typedef typename Mfma<T>::v4 v4;
  v4 cr, ci;
  // S_0 = T_0
  FJob<T> X;  // x_elem's view of the operator (p == 0 form)
  X.op = J.op; X.G = J.G; X.dense = nullptr; X.dsym = nullptr; X.m = J.m; X.n = J.n; X.np = J.np; X.nt = nt;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int a = ti * TB + qi * 16 + Mfma<T>::crow(lane, t), b = tj * TB + qj * 16 + (lane & 15);
    T xr, xi;
    x_elem(X, 0, 0, (const T *)nullptr, a, b, xr, xi);
    cr[t] = xr;
    ci[t] = xi;
  }
  for (int k = 0; k < J.m; ++k) {
    const unsigned tag = (unsigned)k + 1;
    if (k > 0) {  // S_k = T_k - l_k G_{k-1} u_{k-1}, elementwise on this tile
      const int n = J.n;
      const size_t M = (size_t)J.m * n;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int a = ti * TB + qi * 16 + Mfma<T>::crow(lane, t), b = tj * TB + qj * 16 + (lane & 15);
        double vr, vi;
        if (a >= n || b >= n) { vr = (a == b) ? 1.0 : 0.0; vi = 0.0; }
        else {
          const size_t pa = (size_t)k * n + a;
          vr = 0; vi = 0;
          if (a == b) { cd c = ld_c(J.op, FDFD_OP_C * M + pa); vr += c.r; vi += c.i; }
          if (b == (a + 1) % n) { cd c = ld_c(J.op, FDFD_OP_YP * M + pa); vr += c.r; vi += c.i; }
          if (b == (a + n - 1) % n) { cd c = ld_c(J.op, FDFD_OP_YM * M + pa); vr += c.r; vi += c.i; }
          cd sc = cmul(cmul(ld_c(J.op, FDFD_OP_XM * M + pa), cd{(double)cr[t], (double)ci[t]}),
                       ld_c(J.op, FDFD_OP_XP * M + (size_t)(k - 1) * n + b));
          vr -= sc.r; vi -= sc.i;
        }
        cr[t] = (T)vr;
        ci[t] = (T)vi;
      }
    }
    const size_t par = k & 1;
    T *xrow = J.xrow + par * nt * nt * tsz, *xcol = J.xcol + par * nt * nt * tsz;
    T *pinv = J.pinv + par * nt * tsz;
    for (int p = 0; p < nt; ++p) {
      T *prow = xrow + ((size_t)p * nt + tj) * tsz;
      T *pcol = xcol + ((size_t)p * nt + ti) * tsz;
      T *ppiv = pinv + (size_t)p * tsz;
      if (ti == p && tj == p) {
        acc_to_lds<T>(cr, ci, L.Ar, L.Ai, qi, qj, lane);
        __syncthreads();
        tile_inverse<T, PIV>(L, J.status);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int a = qi * 16 + Mfma<T>::crow(lane, t), b = qj * 16 + (lane & 15);
          cr[t] = L.Ar[a][b];
          ci[t] = L.Ai[a][b];
        }
        publish_tile<T>(ppiv, cr, ci, qi, qj, lane);
        stores_complete();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(fpinv + p, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else if (ti == p) {  // X[p, J] <- P X[p, J]
        publish_tile<T>(prow, cr, ci, qi, qj, lane);
        acc_to_lds<T>(cr, ci, L.Ar, L.Ai, qi, qj, lane);
        stores_complete();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(frow + p * nt + tj, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        wait_flag(fpinv + p, tag, abort, J.cap);
        fetch_tile<T>(ppiv, L.Pr, L.Pi);
        __syncthreads();
        cr = v4{0, 0, 0, 0};
        ci = v4{0, 0, 0, 0};
        cgemm_quad<T>(cr, ci, L.Pr, L.Pi, L.Ar, L.Ai, qi, qj, lane, T(1));
      } else if (tj == p) {  // X[I, p] <- -X[I, p] P
        publish_tile<T>(pcol, cr, ci, qi, qj, lane);
        acc_to_lds<T>(cr, ci, L.Br, L.Bi, qi, qj, lane);
        stores_complete();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(fcol + p * nt + ti, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        wait_flag(fpinv + p, tag, abort, J.cap);
        fetch_tile<T>(ppiv, L.Pr, L.Pi);
        __syncthreads();
        cr = v4{0, 0, 0, 0};
        ci = v4{0, 0, 0, 0};
        cgemm_quad<T>(cr, ci, L.Br, L.Bi, L.Pr, L.Pi, qi, qj, lane, T(-1));
      } else {  // X[I, J] -= X[I, p] (P X[p, J])
        wait_flag(fpinv + p, tag, abort, J.cap);
        wait_flag(frow + p * nt + tj, tag, abort, J.cap);
        wait_flag(fcol + p * nt + ti, tag, abort, J.cap);
        {
          const T *const s3[3] = {ppiv, prow, xcol + ((size_t)p * nt + ti) * tsz};
          T (*const d3r[3])[TBP] = {L.Pr, L.Ar, L.Br};
          T (*const d3i[3])[TBP] = {L.Pi, L.Ai, L.Bi};
          fetch_tiles<T, 3>(s3, d3r, d3i);
        }
        __syncthreads();
        v4 rr = {0, 0, 0, 0}, ri = {0, 0, 0, 0};
        cgemm_quad<T>(rr, ri, L.Pr, L.Pi, L.Ar, L.Ai, qi, qj, lane, T(1));
        __syncthreads();  // every wave has read X[p, J]; R takes its place
        acc_to_lds<T>(rr, ri, L.Ar, L.Ai, qi, qj, lane);
        __syncthreads();
        cgemm_quad<T>(cr, ci, L.Br, L.Bi, L.Ar, L.Ai, qi, qj, lane, T(-1));
      }
      __syncthreads();  // LDS is reused by the next step
    }
    // G_k: read by the sweeps after this launch
    T *Gk = J.G + (size_t)k * 2 * np2;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int a = ti * TB + qi * 16 + Mfma<T>::crow(lane, t), b = tj * TB + qj * 16 + (lane & 15);
      Gk[(size_t)a * J.np + b] = cr[t];
      Gk[np2 + (size_t)a * J.np + b] = ci[t];
    }
  }
}

// ---------------------------------------------------------------- sweeps
// Forward, one block: y_k = G_k (r_k - p1(k) y_nb1 - p2(k) y_nb2) for the
// neighbours nb1, nb2 >= 0 (band planes p1, p2; -1: none); y is
// [nrhs][nx][np] complex T. With x (the twist block of a twisted
// factorization, whose y is already its solution) x_k += y_k (FP64).
template <typename T>
__global__ void __launch_bounds__(NT) k_sweep_fwd(const T *__restrict__ G, const double *__restrict__ op,
                                                  const double *__restrict__ r, T *__restrict__ y,
                                                  double *__restrict__ x, int nx, int ny, int np, int k,
                                                  int nrhs, int nb1, int p1, int nb2, int p2) {
  extern __shared__ unsigned char sweep_lds[];
  T *vr = reinterpret_cast<T *>(sweep_lds), *vi = vr + (size_t)MAXR * np;   // [MAXR][np] each
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const size_t np2 = (size_t)np * np, M = (size_t)nx * ny;
  const T *Gk = G + (size_t)k * 2 * np2;
  const int a = blockIdx.x * 4 + w;
  for (int r0 = 0; r0 < nrhs; r0 += MAXR) {
    const int nr = min(MAXR, nrhs - r0);
    for (int e = tid; e < nr * np; e += NT) {
      const int rr = e / np, c = e % np;
      double xr = 0, xi = 0;
      if (c < ny) {
        cd v = ld_c(r, (size_t)(r0 + rr) * M + (size_t)k * ny + c);
        if (nb1 >= 0) {
          const T *yp = y + 2 * (((size_t)(r0 + rr) * nx + nb1) * np + c);
          v = csub(v, cmul(ld_c(op, (size_t)p1 * M + (size_t)k * ny + c), cd{(double)yp[0], (double)yp[1]}));
        }
        if (nb2 >= 0) {
          const T *yp = y + 2 * (((size_t)(r0 + rr) * nx + nb2) * np + c);
          v = csub(v, cmul(ld_c(op, (size_t)p2 * M + (size_t)k * ny + c), cd{(double)yp[0], (double)yp[1]}));
        }
        xr = v.r; xi = v.i;
      }
      vr[rr * np + c] = (T)xr;
      vi[rr * np + c] = (T)xi;
    }
    __syncthreads();
    if (a < np) {
      T ar[MAXR], ai[MAXR];
#pragma unroll
      for (int q = 0; q < MAXR; ++q) { ar[q] = 0; ai[q] = 0; }
      for (int c = lane; c < np; c += 64) {
        const T gr = Gk[(size_t)a * np + c], gi = Gk[np2 + (size_t)a * np + c];
#pragma unroll
        for (int q = 0; q < MAXR; ++q) {
          if (q < nr) {
            ar[q] += gr * vr[q * np + c] - gi * vi[q * np + c];
            ai[q] += gr * vi[q * np + c] + gi * vr[q * np + c];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < MAXR; ++q) {
        for (int off = 32; off >= 1; off >>= 1) {
          ar[q] += __shfl_xor(ar[q], off);
          ai[q] += __shfl_xor(ai[q], off);
        }
      }
      if (lane == 0) {
        for (int q = 0; q < nr; ++q) {
          T *yo = y + 2 * (((size_t)(r0 + q) * nx + k) * np + a);
          yo[0] = ar[q];
          yo[1] = ai[q];
          if (x && a < ny) {
            double *xo = x + 2 * ((size_t)(r0 + q) * M + (size_t)k * ny + a);
            xo[0] += (double)ar[q];
            xo[1] += (double)ai[q];
          }
        }
      }
    }
    __syncthreads();
  }
}

// Backward, one block: d_k = y_k - G_k (p(k) d_nb) in place of y (d_k = y_k
// for nb < 0); x_k += d_k (FP64).
template <typename T>
__global__ void __launch_bounds__(NT) k_sweep_bwd(const T *__restrict__ G, const double *__restrict__ op,
                                                  T *__restrict__ y, double *__restrict__ x,
                                                  int nx, int ny, int np, int k, int nrhs, int nb, int pl) {
  extern __shared__ unsigned char sweep_lds[];
  T *vr = reinterpret_cast<T *>(sweep_lds), *vi = vr + (size_t)MAXR * np;   // [MAXR][np] each
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const size_t np2 = (size_t)np * np, M = (size_t)nx * ny;
  const T *Gk = G + (size_t)k * 2 * np2;
  const int a = blockIdx.x * 4 + w;
  for (int r0 = 0; r0 < nrhs; r0 += MAXR) {
    const int nr = min(MAXR, nrhs - r0);
    for (int e = tid; e < nr * np; e += NT) {
      const int rr = e / np, c = e % np;
      double xr = 0, xi = 0;
      if (c < ny && nb >= 0) {
        const T *dn = y + 2 * (((size_t)(r0 + rr) * nx + nb) * np + c);
        cd v = cmul(ld_c(op, (size_t)pl * M + (size_t)k * ny + c), cd{(double)dn[0], (double)dn[1]});
        xr = v.r; xi = v.i;
      }
      vr[rr * np + c] = (T)xr;
      vi[rr * np + c] = (T)xi;
    }
    __syncthreads();
    if (a < np) {
      T ar[MAXR], ai[MAXR];
#pragma unroll
      for (int q = 0; q < MAXR; ++q) { ar[q] = 0; ai[q] = 0; }
      if (nb >= 0) {
        for (int c = lane; c < np; c += 64) {
          const T gr = Gk[(size_t)a * np + c], gi = Gk[np2 + (size_t)a * np + c];
#pragma unroll
          for (int q = 0; q < MAXR; ++q) {
            if (q < nr) {
              ar[q] += gr * vr[q * np + c] - gi * vi[q * np + c];
              ai[q] += gr * vi[q * np + c] + gi * vr[q * np + c];
            }
          }
        }
#pragma unroll
        for (int q = 0; q < MAXR; ++q) {
          for (int off = 32; off >= 1; off >>= 1) {
            ar[q] += __shfl_xor(ar[q], off);
            ai[q] += __shfl_xor(ai[q], off);
          }
        }
      }
      if (lane == 0) {
        for (int q = 0; q < nr; ++q) {
          T *yk = y + 2 * (((size_t)(r0 + q) * nx + k) * np + a);
          const T dr = yk[0] - ar[q], di = yk[1] - ai[q];
          yk[0] = dr;
          yk[1] = di;
          if (a < ny) {
            double *xo = x + 2 * ((size_t)(r0 + q) * M + (size_t)k * ny + a);
            xo[0] += (double)dr;
            xo[1] += (double)di;
          }
        }
      }
    }
    __syncthreads();
  }
}

// Persistent sweeps: one launch per direction, in two variants that differ
// only in how the next block's vector passes between workgroups (the
// tagged-word sweep for one right-hand side per pass, the flag sweep for
// more). Workgroup w owns rows
// [w*4*SRW, (w+1)*4*SRW) of every block; lane l of a wave covers columns
// l, l+64, ... of its rows. Each workgroup computes its rows of y_k (forward)
// or d_k (backward) and, from them, its rows of the next block's input vector
// (v_{k+1} = r_{k+1} - XM_{k+1} y_k forward, w_{k-1} = XP_{k-1} d_k backward),
// which it publishes; every workgroup then loads the whole vector for the
// next block. Rows of G_k and the inputs of the published entries are loaded
// before the wait, so HBM latency overlaps the synchronization.
// The vector is published as self-validating 64-bit words: 32 bits of the
// value (complex64: one component; FP64: the low or the high half of one)
// and a 32-bit tag naming the step it feeds, written and read with
// agent-scope relaxed atomics (one 64-bit access each, coherent across the
// XCDs' L2 caches). A consumer polls its words until every tag shows the
// step; no flag, no store-completion wait and no second round trip per
// block. Tags grow across the launches of one solve (the host passes each
// launch's base), so a word left by an earlier launch never matches. Two
// buffers alternate by step parity: a workgroup writes a buffer at step s
// only after receiving every entry of step s's input, i.e. after every
// workgroup has finished reading that buffer at step s - 1. A wait that
// exceeds the spin cap raises the abort flag; every later wait then returns
// at once, so the grid always drains. The row sums of a wave
// (SRW rows x RP right-hand sides x re/im) are reduced together. RP, the
// right-hand sides of one pass rounded up to a power of two, is a template
// parameter so the unrolled loops stay small enough for the instruction
// cache; v rows are padded to a multiple of 64 columns with zeros, so the
// inner loop needs no per-lane guard.
constexpr int SRW = 2;              // rows per wave
constexpr int PSWEEP_MAX_N = 768;   // largest padded block of the persistent sweeps
constexpr int MAXCL = PSWEEP_MAX_N / 64;   // columns per lane at that block
constexpr int MAXRHS = 8;           // right-hand sides per pass (complex64)
// right-hand sides per pass: FP64 holds half as many row sums in registers
template <typename T> constexpr int sweep_rhs() { return sizeof(T) == 8 ? MAXRHS / 2 : MAXRHS; }
// tagged words per vector component: the value's 32-bit halves
template <typename T> constexpr int sweep_wpc() { return sizeof(T) == 8 ? 2 : 1; }
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }
inline int pad64(int n) { return (n + 63) / 64 * 64; }

typedef unsigned long long u64;
constexpr int MAXW = PSWEEP_MAX_N / (4 * SRW);   // sweep workgroups at the largest block
constexpr int LL_MAXRHS = 1;                      // right-hand sides of the tagged-word sweep
// Exchange state of one block chain of a persistent sweep
struct SweepChainSync {
  unsigned done[2][MAXW];   // flag sweep, forward / backward: steps finished per workgroup
  // flag sweep: the next block's input vector, [2 parity][RP][np] complex
  double vbuf[2 * MAXRHS * PSWEEP_MAX_N];
  // tagged-word sweep: [2 parity][RP][np][re/im][words per component]
  u64 llbuf[2 * LL_MAXRHS * PSWEEP_MAX_N * 2 * 2];
};
struct SweepSync {
  unsigned abort;
  unsigned pad[63];
  SweepChainSync ch[2];
};

// The block chains of one persistent sweep launch. The factorization is
// either one chain over blocks 0..m-1 or twisted: a top chain 0..c-1, a
// bottom chain m-1..c+1 and the twist block c between them (solved by its
// own launch). Chain i runs on workgroups [i W, (i + 1) W) and visits ns
// blocks k0, k0 + dk, ...; cpl is the band plane of the coupling between a
// block and the chain's previous block, read at the next block
// (forward: v_kn = r_kn - cpl(kn) y_k; backward: w_kn = cpl(kn) d_k). The
// first block's input is r_k0 (forward), zero (backward of a single chain:
// block m-1 has no successor) or cpl(k0) y_t (backward of a twisted chain:
// its coupling to the twist block t).
struct SweepChain {
  int k0, dk, ns, cpl;
  int from_twist;   // backward: the first input is cpl(k0) y_t
  int t;            // the twist block
};
struct SweepChains {
  int nch;
  SweepChain ch[2];
};

// Sums v[0..N) over the 64 lanes of a wave at once: each round halves the
// values a lane holds and exchanges the other half with its partner, so N
// sums cost N-1 + (6 - log2 N) shuffles instead of 6 N. On return every lane
// holds the sum of value (lane >> (6 - log2 N)).
template <int H, typename T, int N>
__device__ inline void multi_sum_round(T (&v)[N], int lane) {
  if constexpr (H >= 1) {
    constexpr int off = 64 * H / N;   // 32 for the first round
    const bool upper = (lane & off) != 0;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const T send = upper ? v[j] : v[j + H];
      const T keep = upper ? v[j + H] : v[j];
      v[j] = keep + __shfl_xor(send, off);
    }
    multi_sum_round<H / 2>(v, lane);
  }
}
template <typename T, int N>
__device__ inline T wave_multi_sum(T (&v)[N], int lane) {
  static_assert(N >= 2 && N <= 64 && (N & (N - 1)) == 0, "N must be a power of two <= 64");
  multi_sum_round<N / 2>(v, lane);
  T t = v[0];
#pragma unroll
  for (int off = 32 / N; off >= 1; off >>= 1) t += __shfl_xor(t, off);
  return t;
}

// publish one component of a vector entry with its tag
__device__ inline void publish_tagged(u64 *w, float v, unsigned tag) {
  st_xcd(w, (u64)__float_as_uint(v) | ((u64)tag << 32));
}
__device__ inline void publish_tagged(u64 *w, double v, unsigned tag) {
  const u64 bits = (u64)__double_as_longlong(v);
  st_xcd(w, (bits & 0xffffffffull) | ((u64)tag << 32));
  st_xcd(w + 1, (bits >> 32) | ((u64)tag << 32));
}

// a complex entry of a published vector, one load per entry
__device__ inline void ld_xcd_c(const float *p, float &re, float &im) {
  const unsigned long long v = ld_xcd(reinterpret_cast<const unsigned long long *>(p));
  re = __uint_as_float((unsigned)(v & 0xffffffffull));
  im = __uint_as_float((unsigned)(v >> 32));
}
__device__ inline void ld_xcd_c(const double *p, double &re, double &im) {
  re = ld_xcd(p);
  im = ld_xcd(p + 1);
}

// The flag sweep (two or more right-hand sides per pass): each workgroup
// publishes its entries of the next block's vector with agent-scope relaxed
// atomics, waits for its own stores to complete, passes a barrier and raises
// its per-workgroup flag; consumers wait for all W flags of the previous
// step, then load the vector. Per block that is three memory round trips,
// but the vector moves as plain 8-byte values: with several right-hand sides
// this is faster than the tagged words below, which double the bytes and
// re-poll every word.
template <typename T, int RP>
__global__ void __launch_bounds__(NT) k_psweep(const T *__restrict__ G, const double *__restrict__ op,
                                               const double *__restrict__ r, T *y, double *__restrict__ x,
                                               int nx, int ny, int np, int nrhs, int dir, SweepChains chs,
                                               SweepSync *sy, unsigned cap) {
  extern __shared__ unsigned char smem_raw[];
  const int nv = (np + 63) / 64 * 64;                  // padded row length of v
  T *vr = reinterpret_cast<T *>(smem_raw);             // [RP][nv]
  T *vi = vr + (size_t)RP * nv;
  const int W = gridDim.x / chs.nch, grp = blockIdx.x / W, wl = blockIdx.x % W;
  const SweepChain ch = chs.ch[grp];
  SweepChainSync *cs = &sy->ch[grp];
  T *vb = reinterpret_cast<T *>(cs->vbuf);             // [2][RP][np][2]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const size_t np2 = (size_t)np * np, M = (size_t)nx * ny;
  const int row0 = (wl * 4 + wv) * SRW;
  // the (row, rhs, re/im) value this lane finishes after the reduction
  static_assert(RP >= 1 && RP <= MAXRHS && (RP & (RP - 1)) == 0, "RP must be a power of two");
  constexpr int NV = 2 * SRW * RP, SH = 6 - ilog2(NV);
  const int jv = lane >> SH;   // lanes jv << SH ... all hold value jv
  const int my_ri = jv & 1, my_q = (jv >> 1) % RP, my_rr = jv / (2 * RP);
  const int my_a = row0 + my_rr;
  const bool my_store = (lane & ((1 << SH) - 1)) == 0 && my_q < nrhs && my_a < np;
  const bool my_real = my_q < nrhs && my_a < ny;
  const int cpl_plane = ch.cpl;
  for (int s = 0; s < ch.ns; ++s) {
    const int k = ch.k0 + ch.dk * s;
    const T *Gk = G + (size_t)k * 2 * np2;
    // the first backward block of a single chain has no successor: d = y
    const bool active = dir == 0 || s > 0 || ch.from_twist;
    const int kn = k + ch.dk;                            // the next block
    const bool has_next = s + 1 < ch.ns;
    // 1. this workgroup's rows of G_k and the inputs of its published
    //    entries (independent of the previous block)
    T gr[SRW][MAXCL], gi[SRW][MAXCL];
#pragma unroll
    for (int rr = 0; rr < SRW; ++rr) {
      const int a = row0 + rr;
#pragma unroll
      for (int cc = 0; cc < MAXCL; ++cc) {
        const int c = lane + 64 * cc;
        const bool ok = a < np && c < np && active;
        gr[rr][cc] = ok ? Gk[(size_t)a * np + c] : T(0);
        gi[rr][cc] = ok ? Gk[np2 + (size_t)a * np + c] : T(0);
      }
    }
    cd rn = {0, 0}, cn = {0, 0};   // forward: r_kn, cpl(kn); backward: cpl(kn) (this lane's row)
    if (my_real && has_next) {
      if (dir == 0) rn = ld_c(r, (size_t)my_q * M + (size_t)kn * ny + my_a);
      cn = ld_c(op, cpl_plane * M + (size_t)kn * ny + my_a);
    }
    T yown = 0;   // backward: this lane's entry of y_k, written by the forward launch
    T *my_y = y + 2 * (((size_t)my_q * nx + k) * np + my_a) + my_ri;
    if (dir == 1 && my_store) yown = *my_y;
    // 2. this block's input vector into LDS: the chain's first input at its
    //    first step, else the vector all workgroups published at the previous
    //    step (every load issued before any use)
    if (s == 0) {
      constexpr int CU = (PSWEEP_MAX_N + NT - 1) / NT;
      double ar[RP][CU], ai[RP][CU];
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int c = tid + u * NT;
          const bool ok = q < nrhs && c < ny;
          cd v = {0, 0};
          if (ok && dir == 0) {
            v = ld_c(r, (size_t)q * M + (size_t)k * ny + c);
          } else if (ok && ch.from_twist) {
            const T *yt = y + 2 * (((size_t)q * nx + ch.t) * np + c);
            v = cmul(ld_c(op, (size_t)cpl_plane * M + (size_t)k * ny + c), cd{(double)yt[0], (double)yt[1]});
          }
          ar[q][u] = v.r;
          ai[q][u] = v.i;
        }
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int c = tid + u * NT;
          if (c < nv) {
            vr[q * nv + c] = (T)ar[q][u];
            vi[q * nv + c] = (T)ai[q][u];
          }
        }
      __syncthreads();
    } else {
      wait_all_flags(cs->done[dir], W, (unsigned)s, &sy->abort, cap);
      const T *vin = vb + (size_t)(s & 1) * RP * np * 2;
      constexpr int CU = (PSWEEP_MAX_N + NT - 1) / NT;
      T br[RP][CU], bi[RP][CU];
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int c = tid + u * NT;
          br[q][u] = T(0);
          bi[q][u] = T(0);
          if (q < nrhs && c < np) ld_xcd_c(vin + ((size_t)q * np + c) * 2, br[q][u], bi[q][u]);
        }
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int c = tid + u * NT;
          if (c < nv) {
            vr[q * nv + c] = br[q][u];
            vi[q * nv + c] = bi[q][u];
          }
        }
      __syncthreads();
    }
    // 3. partial row dot products for every (row, rhs) pair of this wave
    T acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0;
    if (active) {
#pragma unroll
      for (int cc = 0; cc < MAXCL; ++cc) {
        if (64 * cc < np) {   // uniform; columns c >= np hold G = 0 and v = 0
          const int c = lane + 64 * cc;
#pragma unroll
          for (int q = 0; q < RP; ++q) {
            const T br = vr[q * nv + c], bi = vi[q * nv + c];
#pragma unroll
            for (int rr = 0; rr < SRW; ++rr) {
              T &sr = acc[(rr * RP + q) * 2], &si = acc[(rr * RP + q) * 2 + 1];
              sr += gr[rr][cc] * br - gi[rr][cc] * bi;
              si += gr[rr][cc] * bi + gi[rr][cc] * br;
            }
          }
        }
      }
    }
    const T sum = wave_multi_sum(acc, lane);
    // 4. this lane's entry of y_k (forward) or d_k (backward)
    const T out = dir == 0 ? sum : yown - sum;
    if (my_store) {
      if (dir == 0) *my_y = out;   // read by the backward launch
      else if (my_a < ny) x[2 * ((size_t)my_q * M + (size_t)k * ny + my_a) + my_ri] += (double)out;
    }
    // 5. publish the next block's input entry for this row:
    //    forward v_kn = r_kn - cpl(kn) y_k, backward w_kn = cpl(kn) d_k
    const T other = __shfl_xor(out, 1 << SH);   // the other component of this (row, rhs)
    if (my_store && has_next) {
      const double ore = (double)(my_ri == 0 ? out : other), oim = (double)(my_ri == 0 ? other : out);
      const cd t = cmul(cn, cd{ore, oim});
      const cd v = dir == 0 ? csub(rn, t) : t;
      T *vout = vb + (size_t)((s + 1) & 1) * RP * np * 2;
      st_xcd(vout + ((size_t)my_q * np + my_a) * 2 + my_ri, (T)(my_ri == 0 ? v.r : v.i));
    }
    // 6. publish: every wave's stores complete, then this workgroup's flag
    stores_complete();
    __syncthreads();
    if (tid == 0)
      __hip_atomic_store(&cs->done[dir][wl], (unsigned)(s + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}


template <typename T, int RP>
__global__ void __launch_bounds__(NT) k_psweep_ll(const T *__restrict__ G, const double *__restrict__ op,
                                                  const double *__restrict__ r, T *y, double *__restrict__ x,
                                                  int nx, int ny, int np, int nrhs, int dir, SweepChains chs,
                                                  SweepSync *sy, unsigned cap, unsigned tag0) {
  extern __shared__ unsigned char smem_raw[];
  const int nv = (np + 63) / 64 * 64;                  // padded row length of v
  T *vr = reinterpret_cast<T *>(smem_raw);             // [RP][nv]
  T *vi = vr + (size_t)RP * nv;
  const int W = gridDim.x / chs.nch, grp = blockIdx.x / W, wl = blockIdx.x % W;
  const SweepChain ch = chs.ch[grp];
  SweepChainSync *cs = &sy->ch[grp];
  constexpr int WPC = sweep_wpc<T>();
  const size_t vstride = (size_t)RP * np * 2 * WPC;    // words of one parity buffer
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const size_t np2 = (size_t)np * np, M = (size_t)nx * ny;
  const int row0 = (wl * 4 + wv) * SRW;
  // the (row, rhs, re/im) value this lane finishes after the reduction
  static_assert(RP >= 1 && RP <= MAXRHS && (RP & (RP - 1)) == 0, "RP must be a power of two");
  constexpr int NV = 2 * SRW * RP, SH = 6 - ilog2(NV);
  const int jv = lane >> SH;   // lanes jv << SH ... all hold value jv
  const int my_ri = jv & 1, my_q = (jv >> 1) % RP, my_rr = jv / (2 * RP);
  const int my_a = row0 + my_rr;
  const bool my_store = (lane & ((1 << SH) - 1)) == 0 && my_q < nrhs && my_a < np;
  const bool my_real = my_q < nrhs && my_a < ny;
  const int cpl_plane = ch.cpl;
  // words this thread receives per step: component c of rhs q is word
  // ((q * np + c) * 2 + re/im) * WPC + half
  constexpr int CW = (RP * PSWEEP_MAX_N * 2 * WPC + NT - 1) / NT;
  constexpr int CH = CW < 24 ? CW : 24;
  static_assert(CW <= 64, "receive mask");
  const int nword = nrhs * np * 2 * WPC;
  // word tid + u * NT is component (re/im) w_ri, half w_half of entry c of
  // rhs q; consecutive words of a thread are WSTEP entries apart
  constexpr int WSTEP = NT / (2 * WPC);
  const int w_half = tid % WPC, w_ri = (tid / WPC) & 1, w_qc = tid / (2 * WPC);
  const int w_q0 = w_qc / np, w_c0 = w_qc % np;
  T *const w_dst = w_ri == 0 ? vr : vi;
  u64 expect = 0;
#pragma unroll
  for (int u = 0; u < CW; ++u)
    if (tid + u * NT < nword) expect |= 1ull << u;
  // rows of unused right-hand sides and the column padding stay zero
  for (int e = tid; e < RP * nv; e += NT) {
    vr[e] = T(0);
    vi[e] = T(0);
  }
  for (int s = 0; s < ch.ns; ++s) {
    const int k = ch.k0 + ch.dk * s;
    const T *Gk = G + (size_t)k * 2 * np2;
    // the first backward block of a single chain has no successor: d = y
    const bool active = dir == 0 || s > 0 || ch.from_twist;
    const int kn = k + ch.dk;                            // the next block
    const bool has_next = s + 1 < ch.ns;
    // 1. this workgroup's rows of G_k and the inputs of its published
    //    entries (independent of the previous block)
    T gr[SRW][MAXCL], gi[SRW][MAXCL];
#pragma unroll
    for (int rr = 0; rr < SRW; ++rr) {
      const int a = row0 + rr;
#pragma unroll
      for (int cc = 0; cc < MAXCL; ++cc) {
        const int c = lane + 64 * cc;
        const bool ok = a < np && c < np && active;
        gr[rr][cc] = ok ? Gk[(size_t)a * np + c] : T(0);
        gi[rr][cc] = ok ? Gk[np2 + (size_t)a * np + c] : T(0);
      }
    }
    cd rn = {0, 0}, cn = {0, 0};   // forward: r_kn, cpl(kn); backward: cpl(kn) (this lane's row)
    if (my_real && has_next) {
      if (dir == 0) rn = ld_c(r, (size_t)my_q * M + (size_t)kn * ny + my_a);
      cn = ld_c(op, cpl_plane * M + (size_t)kn * ny + my_a);
    }
    T yown = 0;   // backward: this lane's entry of y_k, written by the forward launch
    T *my_y = y + 2 * (((size_t)my_q * nx + k) * np + my_a) + my_ri;
    if (dir == 1 && my_store) yown = *my_y;
    // 2. this block's input vector into LDS: the chain's first input at its
    //    first step, else the vector all workgroups published at the previous
    //    step
    __syncthreads();   // every wave has finished reading the previous block's vector
    if (s == 0) {
      constexpr int CU = (PSWEEP_MAX_N + NT - 1) / NT;
      double ar[RP][CU], ai[RP][CU];
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int c = tid + u * NT;
          const bool ok = q < nrhs && c < ny;
          cd v = {0, 0};
          if (ok && dir == 0) {
            v = ld_c(r, (size_t)q * M + (size_t)k * ny + c);
          } else if (ok && ch.from_twist) {
            const T *yt = y + 2 * (((size_t)q * nx + ch.t) * np + c);
            v = cmul(ld_c(op, (size_t)cpl_plane * M + (size_t)k * ny + c), cd{(double)yt[0], (double)yt[1]});
          }
          ar[q][u] = v.r;
          ai[q][u] = v.i;
        }
#pragma unroll
      for (int q = 0; q < RP; ++q)
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int c = tid + u * NT;
          if (c < nv) {
            vr[q * nv + c] = (T)ar[q][u];
            vi[q * nv + c] = (T)ai[q][u];
          }
        }
    } else {
      const u64 *vin = cs->llbuf + (size_t)(s & 1) * vstride;
      const unsigned tag = tag0 + (unsigned)s;
      unsigned spins = 0;
      bool give_up = false;
      int qa = w_q0, ca = w_c0;   // (rhs, entry) of the chunk's first word
      // words in chunks of at most CH loads in flight (bounds the registers)
#pragma unroll
      for (int u0 = 0; u0 < CW; u0 += CH) {
        u64 pending = (expect >> u0) & ((CH >= 64) ? ~0ull : ((1ull << CH) - 1));
        // many words: wait on one of them first, so the spin reads one word
        // per thread instead of all of them
        if (__popcll(pending) > 16) {
          const u64 *first = vin + tid + (u0 + __builtin_ctzll(pending)) * NT;
          while (!give_up && (unsigned)(ld_xcd(first) >> 32) != tag) {
            if ((++spins & 63) == 0 && ld_xcd(&sy->abort)) give_up = true;
            if (spins > cap) {
              st_xcd(&sy->abort, 1u);
              give_up = true;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        while (pending && !give_up) {
          u64 w[CH];
#pragma unroll
          for (int u = 0; u < CH; ++u)
            w[u] = (u0 + u < CW && ((pending >> u) & 1)) ? ld_xcd(vin + tid + (u0 + u) * NT) : 0ull;
          int q = qa, c = ca;
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            if (u0 + u < CW && ((pending >> u) & 1) && (unsigned)(w[u] >> 32) == tag) {
              pending &= ~(1ull << u);
              const unsigned bits = (unsigned)(w[u] & 0xffffffffull);
              __builtin_memcpy(reinterpret_cast<unsigned char *>(w_dst + q * nv + c) + 4 * w_half, &bits, 4);
            }
            c += WSTEP;
            while (c >= np) { c -= np; ++q; }
          }
          if (!pending) break;
          if ((++spins & 63) == 0 && ld_xcd(&sy->abort)) give_up = true;
          if (spins > cap) {
            st_xcd(&sy->abort, 1u);
            give_up = true;
          }
          __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {   // advance to the next chunk's first word
          ca += WSTEP;
          while (ca >= np) { ca -= np; ++qa; }
        }
      }
    }
    __syncthreads();
    // 3. partial row dot products for every (row, rhs) pair of this wave
    T acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0;
    if (active) {
#pragma unroll
      for (int cc = 0; cc < MAXCL; ++cc) {
        if (64 * cc < np) {   // uniform; columns c >= np hold G = 0 and v = 0
          const int c = lane + 64 * cc;
#pragma unroll
          for (int q = 0; q < RP; ++q) {
            const T br = vr[q * nv + c], bi = vi[q * nv + c];
#pragma unroll
            for (int rr = 0; rr < SRW; ++rr) {
              T &sr = acc[(rr * RP + q) * 2], &si = acc[(rr * RP + q) * 2 + 1];
              sr += gr[rr][cc] * br - gi[rr][cc] * bi;
              si += gr[rr][cc] * bi + gi[rr][cc] * br;
            }
          }
        }
      }
    }
    const T sum = wave_multi_sum(acc, lane);
    // 4. this lane's entry of y_k (forward) or d_k (backward)
    const T out = dir == 0 ? sum : yown - sum;
    if (my_store) {
      if (dir == 0) *my_y = out;   // read by the backward launch
      else if (my_a < ny) x[2 * ((size_t)my_q * M + (size_t)k * ny + my_a) + my_ri] += (double)out;
    }
    // 5. publish the next block's input entry for this row:
    //    forward v_kn = r_kn - cpl(kn) y_k, backward w_kn = cpl(kn) d_k
    const T other = __shfl_xor(out, 1 << SH);   // the other component of this (row, rhs)
    if (my_store && has_next) {
      const double ore = (double)(my_ri == 0 ? out : other), oim = (double)(my_ri == 0 ? other : out);
      const cd t = cmul(cn, cd{ore, oim});
      const cd v = dir == 0 ? csub(rn, t) : t;
      u64 *vout = cs->llbuf + (size_t)((s + 1) & 1) * vstride;
      publish_tagged(vout + (((size_t)my_q * np + my_a) * 2 + my_ri) * WPC, (T)(my_ri == 0 ? v.r : v.i),
                     tag0 + (unsigned)(s + 1));
    }
  }
}

// ------------------------------------------------------- x-wrap correction
// The block sweep factors T, the operator without the two x-wrap couplings
// E = U C V^T: block 0 couples to block m-1 through XM of block 0, and block
// m-1 to block 0 through XP of block m-1 (U = [E_0, E_{m-1}],
// C = diag(XM_0, XP_{m-1}), V = [E_{m-1}, E_0]). Woodbury gives
//   A^-1 = T^-1 - T^-1 U C K^-1 V^T T^-1,  K = I + V^T T^-1 U C,
// where V^T T^-1 U = [[T^-1_{m-1,0}, T^-1_{m-1,m-1}], [T^-1_{0,0}, T^-1_{0,m-1}]].
// T^-1_{m-1,m-1} = G_{m-1}. The other corner blocks come from two backward
// chains after the factorization (Sigma is the selected-inversion recurrence
// for the diagonal blocks of T^-1, Z the last block column):
//   Sigma_k = G_k + G_k XP_k Sigma_{k+1} XM_{k+1} G_k,   Z_k = -G_k XP_k Z_{k+1},
// both starting from G_{m-1}; T^-1_{m-1,0} follows from T^-1_{0,m-1} by the
// stretch symmetrizer D (D T is complex symmetric, so
// (T^-1)_{ji} = D_i (T^-1)_{ij} / D_j). K (2n x 2n, complex FP64) is inverted
// once per factorization by Gauss-Jordan with partial pivoting and kept after
// the block inverses.
// With PMLs of at least WRAP_MIN_PML cells on both axes the far corners
// T^-1_{0,m-1} and T^-1_{m-1,0} (waves crossing the whole domain and both
// x-PMLs) are dropped and the Sigma chain starts at block
// L = NPML_x + WRAP_EXTRA, past the x-PML, with Sigma_L = G_L. What the
// truncation leaves out is the part of T^-1_{0,0} reflected from beyond L,
// which has crossed the x-PML twice and does not shrink with a longer chain;
// the refinement takes as many steps with L = NPML_x + 4 as with
// 2 NPML_x + 16. The chains shrink from 3 (m - 1) products to 2 L. Without
// such PMLs (periodic axes, where waves along the blocks are never absorbed)
// the chains run over all blocks and K is exact.

// C = D + sign * A * (rs . B . cs) on n_p x n_p planar matrices, 32 x 32 output
// tiles per workgroup, MFMA complex products (rs / cs: optional complex row /
// column scalings of B from band planes; entries beyond n scale by zero).
template <typename T> struct GemmJob {
  const T *A, *B, *D;
  T *C;
  const double *rs, *cs;
  T sign;
  // [4][np] D, 1/D of the block symmetrizer: C is D-symmetric
  // (C[y][x] = (D[x] / D[y]) C[x][y]); only its tiles I <= J are computed and
  // the others written as their mirror images. Null: every tile.
  const T *dsym;
};
template <typename T> struct GemmBatch {
  int njobs;
  GemmJob<T> job[2];
};

// k tile kt of the operands of output tile (ti, tj): A[ti, kt] and the
// scaled B[kt, tj]; thread t holds four consecutive elements of row t/8 of
// each (one 16-byte load per plane; zeros past nt)
template <typename T>
__device__ __forceinline__ void gemm_fetch(const GemmJob<T> &J, int n, int np, int nt, int ti, int tj, int kt,
                                           int tid, T (&ar)[TB * TB / NT], T (&ai)[TB * TB / NT],
                                           T (&br)[TB * TB / NT], T (&bi)[TB * TB / NT]) {
  static_assert(TB * TB / NT == 4, "four elements per thread");
  const size_t np2 = (size_t)np * np;
  if (kt >= nt) {
#pragma unroll
    for (int u = 0; u < 4; ++u) { ar[u] = ai[u] = br[u] = bi[u] = T(0); }
    return;
  }
  const int a = tid >> 3, b0 = (tid & 7) * 4;
  const int brow = kt * TB + a;
  const size_t ia = (size_t)(ti * TB + a) * np + kt * TB + b0, ib = (size_t)brow * np + tj * TB + b0;
  ld_row4(J.A + ia, ar);
  ld_row4(J.A + np2 + ia, ai);
  T xr[4], xi[4];
  ld_row4(J.B + ib, xr);
  ld_row4(J.B + np2 + ib, xi);
  const cd rsc = J.rs ? (brow < n ? ld_c(J.rs, brow) : cd{0, 0}) : cd{1, 0};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    double yr = xr[u], yi = xi[u];
    if (J.rs) {
      const double zr = rsc.r * yr - rsc.i * yi, zi = rsc.r * yi + rsc.i * yr;
      yr = zr; yi = zi;
    }
    if (J.cs) {
      const int bcol = tj * TB + b0 + u;
      const cd sc = bcol < n ? ld_c(J.cs, bcol) : cd{0, 0};
      const double zr = sc.r * yr - sc.i * yi, zi = sc.r * yi + sc.i * yr;
      yr = zr; yi = zi;
    }
    br[u] = (T)yr;
    bi[u] = (T)yi;
  }
}

// Output tile (ti, tj) of job jb per workgroup of GT threads: two groups of
// four waves take alternate k tiles; each group keeps the loads of its next
// tile in flight while it multiplies the current one, and group 1's partial
// sums are added to group 0's in a fixed order at the end.
constexpr int GT = 2 * NT;
template <typename T>
__global__ void __launch_bounds__(GT) k_cgemm(GemmBatch<T> batch, int n, int np) {
  __shared__ T Ar[2][TB][TBP], Ai[2][TB][TBP], Br[2][TB][TBP], Bi[2][TB][TBP];
  const int nt = np / TB;
  // both jobs of a launch compute the same tiles: all of them, or I <= J
  const int per = batch.job[0].dsym ? nt * (nt + 1) / 2 : nt * nt;
  const int jb = blockIdx.x / per, t = blockIdx.x % per;
  const GemmJob<T> J = jb == 0 ? batch.job[0] : batch.job[1];
  const int tu = J.dsym ? upper_tile(nt, t) : t;
  const int ti = tu / nt, tj = tu % nt;
  const int grp = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int lane = tid & 63, wv = tid >> 6, qi = wv >> 1, qj = wv & 1;
  const size_t np2 = (size_t)np * np;
  constexpr int PER = TB * TB / NT;
  T ar[PER], ai[PER], br[PER], bi[PER];
  typename Mfma<T>::v4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  const int iters = (nt + 1) / 2;
  gemm_fetch<T>(J, n, np, nt, ti, tj, grp, tid, ar, ai, br, bi);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {   // gemm_fetch's mapping: row tid/8, columns 4 (tid % 8) + u
      const int a = tid >> 3, b = (tid & 7) * 4 + u;
      Ar[grp][a][b] = ar[u];
      Ai[grp][a][b] = ai[u];
      Br[grp][a][b] = br[u];
      Bi[grp][a][b] = bi[u];
    }
    __syncthreads();
    const int kt = 2 * it + grp;
    if (it + 1 < iters) gemm_fetch<T>(J, n, np, nt, ti, tj, kt + 2, tid, ar, ai, br, bi);
    if (kt < nt) cgemm_quad<T>(cr, ci, Ar[grp], Ai[grp], Br[grp], Bi[grp], qi, qj, lane, J.sign);
    __syncthreads();
  }
  if (grp == 1) acc_to_lds<T>(cr, ci, Ar[0], Ai[0], qi, qj, lane);
  __syncthreads();
  if (grp == 1) return;
  const bool mirror = J.dsym && ti != tj;
  // the mirror tile (J, I): this lane's element (a, b) goes to row
  // y = 32 J + b, column x = 32 I + a, scaled by D[x] / D[y]
  T dxr[4] = {0, 0, 0, 0}, dxi[4] = {0, 0, 0, 0}, eyr = 0, eyi = 0;
  if (mirror) {
    const int y = tj * TB + qj * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int x = ti * TB + qi * 16 + Mfma<T>::crow(lane, q);
      dxr[q] = J.dsym[x];
      dxi[q] = J.dsym[np + x];
    }
    eyr = J.dsym[2 * np + y];
    eyi = J.dsym[3 * np + y];
  }
  T mr[4], mi[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int a = qi * 16 + Mfma<T>::crow(lane, q), b = qj * 16 + (lane & 15);
    const int row = ti * TB + a, col = tj * TB + b;
    const size_t idx = (size_t)row * np + col;
    T vr = cr[q] + Ar[0][a][b], vi = ci[q] + Ai[0][a][b];
    if (J.D) { vr += J.D[idx]; vi += J.D[np2 + idx]; }
    J.C[idx] = vr;
    J.C[np2 + idx] = vi;
    const T rr = dxr[q] * eyr - dxi[q] * eyi, ri = dxr[q] * eyi + dxi[q] * eyr;
    mr[q] = rr * vr - ri * vi;
    mi[q] = rr * vi + ri * vr;
  }
  if (mirror) {
    const int y = tj * TB + qj * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const size_t midx = (size_t)y * np + ti * TB + qi * 16 + Mfma<T>::crow(lane, q);
      J.C[midx] = mr[q];
      J.C[np2 + midx] = mi[q];
    }
  }
}

__device__ inline cd stretch_d(const fdfd_grid &g, int i, int j) {
  const double hx = ((double)g.nx * g.dl) / g.nx, hy = ((double)g.ny * g.dl) / g.ny;
  return cmul(sfactor(0, i, g.nx, g.npml_x, hx, g.omega, g.L0), sfactor(0, j, g.ny, g.npml_y, hy, g.omega, g.L0));
}

// K = I + V^T T^-1 U C as a row-major complex FP64 matrix of order 2n
constexpr int WRAP_MIN_PML = 4, WRAP_EXTRA = 4;

// Y = NULL: the far corners are taken as zero
template <typename T>
__global__ void __launch_bounds__(NT) k_wrap_capacitance(fdfd_grid g, const double *__restrict__ op,
                                                         const T *__restrict__ Glast, const T *__restrict__ Sig,
                                                         const T *__restrict__ Y, int np, double *__restrict__ K) {
  const int n = g.ny, m = g.nx, N = 2 * n;
  const int j = blockIdx.x * NT + threadIdx.x, i = blockIdx.y;
  if (j >= N) return;
  const size_t np2 = (size_t)np * np, M = (size_t)m * n;
  const int bi = i / n, bj = j / n, a = i % n, b = j % n;
  cd v = {0, 0};
  if (bi == bj && !Y) {
    // far corner dropped
  } else if (bi == 0 && bj == 0) {   // T^-1_{m-1,0}[a][b] = D_{0,b} T^-1_{0,m-1}[b][a] / D_{m-1,a}
    const size_t e = (size_t)b * np + a;
    v = cmul(cd{(double)Y[e], (double)Y[np2 + e]}, cmul(stretch_d(g, 0, b), cinv(stretch_d(g, m - 1, a))));
  } else {
    const T *src = bi == 0 ? Glast : (bj == 0 ? Sig : Y);
    const size_t e = (size_t)a * np + b;
    v = cd{(double)src[e], (double)src[np2 + e]};
  }
  const cd c = bj == 0 ? ld_c(op, FDFD_OP_XM * M + b) : ld_c(op, FDFD_OP_XP * M + (size_t)(m - 1) * n + b);
  v = cmul(v, c);
  if (i == j) v.r += 1.0;
  st_c(K, (size_t)i * N + j, v);
}

// Gauss-Jordan inverse of K with partial pivoting, one launch per column j:
// the launch reads the matrix before step j from one buffer and writes the
// matrix after it to the other, so no element is read after it is written.
// Workgroup (bx, by) owns rows [KGJ_ROWS by, +KGJ_ROWS) x columns
// [256 bx, +256) and repeats the pivot search over column j (identical
// result everywhere), reading it from a contiguous copy (colin) that the
// previous launch wrote alongside column j of its output (colout gets
// column j + 1 of this launch's output). Row interchanges are applied as the
// rows are written; the column swaps are undone at the end
// (k_kgj_unpermute).
constexpr int KGJ_ROWS = 32;
__global__ void __launch_bounds__(NT) k_kgj_col(const double *__restrict__ Ks, double *__restrict__ Kd, int N,
                                                int j, const double *__restrict__ colin, double *__restrict__ colout,
                                                int *piv, int *status) {
  __shared__ double bv[NT];
  __shared__ int bi[NT];
  const int tid = threadIdx.x;
  double best = -1;
  int arg = j;
  for (int i = j + tid; i < N; i += NT) {
    const cd z = ld_c(colin, i);
    const double v = fabs(z.r) + fabs(z.i);
    if (v > best) { best = v; arg = i; }
  }
  bv[tid] = best;
  bi[tid] = arg;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (tid < st) {
      const double o = bv[tid + st];
      const int oi = bi[tid + st];
      if (o > bv[tid] || (o == bv[tid] && oi < bi[tid])) { bv[tid] = o; bi[tid] = oi; }
    }
    __syncthreads();
  }
  const int p = bi[0];
  const cd pv = ld_c(colin, p);
  const cd d = (pv.r == 0 && pv.i == 0) ? cd{0, 0} : cinv(pv);
  if (blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) {
    piv[j] = p;
    if (!(bv[0] > 0) || !isfinite(bv[0])) status[0] = 1;
  }
  const int c = blockIdx.x * NT + tid;
  if (c >= N) return;
  const cd h = ld_c(Ks, (size_t)p * N + c);   // pivot row (row p before the interchange)
  for (int r = 0; r < KGJ_ROWS; ++r) {
    const int i = blockIdx.y * KGJ_ROWS + r;
    if (i >= N) break;
    cd v;
    if (i == j) {
      v = c == j ? d : cmul(h, d);
    } else {
      const int sr = i == p ? j : i;   // row p now holds the old row j
      const cd f = ld_c(colin, sr);
      const cd g = cmul(f, d);
      v = c == j ? cd{-g.r, -g.i} : csub(ld_c(Ks, (size_t)sr * N + c), cmul(g, h));
    }
    st_c(Kd, (size_t)i * N + c, v);
    if (c == j + 1) st_c(colout, i, v);
  }
}

// column 0 of K into the pivot-search copy
__global__ void __launch_bounds__(NT) k_kgj_col0(const double *__restrict__ K, int N, double *__restrict__ col) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < N) st_c(col, i, ld_c(K, (size_t)i * N));
}

__global__ void __launch_bounds__(NT) k_kgj_unpermute(double *K, int N, const int *__restrict__ piv) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= N) return;
  for (int j = N - 1; j >= 0; --j) {
    const int p = piv[j];
    if (p != j) {
      const cd x = ld_c(K, (size_t)i * N + j), y = ld_c(K, (size_t)i * N + p);
      st_c(K, (size_t)i * N + j, y);
      st_c(K, (size_t)i * N + p, x);
    }
  }
}

// f = -U C K^-1 V^T t for every right-hand side: one wave per row of K^-1;
// rows < n land in block 0 (times XM_0), the others in block m-1 (times
// XP_{m-1}). f must be zero elsewhere.
__global__ void __launch_bounds__(NT) k_wrap_rhs(fdfd_grid g, const double *__restrict__ op,
                                                 const double *__restrict__ Kinv, const double *__restrict__ t,
                                                 double *__restrict__ f) {
  const int n = g.ny, m = g.nx, N = 2 * n;
  const int lane = threadIdx.x & 63, i = blockIdx.x * (NT / 64) + (threadIdx.x >> 6), q = blockIdx.y;
  if (i >= N) return;
  const size_t M = (size_t)m * n;
  const double *tq = t + 2 * M * q;
  double sr = 0, si = 0;
  for (int j = lane; j < N; j += 64) {
    const cd kv = ld_c(Kinv, (size_t)i * N + j);
    const cd sv = j < n ? ld_c(tq, (size_t)(m - 1) * n + j) : ld_c(tq, (size_t)(j - n));
    sr += kv.r * sv.r - kv.i * sv.i;
    si += kv.r * sv.i + kv.i * sv.r;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    sr += __shfl_xor(sr, off);
    si += __shfl_xor(si, off);
  }
  if (lane == 0) {
    const size_t pos = i < n ? (size_t)i : (size_t)(m - 1) * n + (i - n);
    const cd c = i < n ? ld_c(op, FDFD_OP_XM * M + i) : ld_c(op, FDFD_OP_XP * M + (size_t)(m - 1) * n + (i - n));
    const cd w = cmul(c, cd{sr, si});
    st_c(f + 2 * M * q, pos, cd{-w.r, -w.i});
  }
}

__global__ void __launch_bounds__(NT) k_axpy(size_t len, const double *__restrict__ a, double *__restrict__ y) {
  for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < len; e += (size_t)gridDim.x * NT) y[e] += a[e];
}

// Truncated wrap (far corners dropped): K = [[I, B], [A, I]] with
// B = G_{m-1} C2, A = Sigma_0 C1, so only S = I - A B (order n) is inverted,
// and K^-1 [s1; s2] is w2 = S^-1 (s2 - A s1), w1 = s1 - B w2.
// S from its negative product -A B (planar T, stride np): planar complex
// FP64 ([2][np][np], identity on the padding) when tiled, else row-major
// complex FP64 of order n.
template <typename T>
__global__ void __launch_bounds__(NT) k_wrap_schur(const T *__restrict__ negAB, int np, int n,
                                                   double *__restrict__ S, int tiled) {
  const int j = blockIdx.x * NT + threadIdx.x, i = blockIdx.y;
  const size_t e = (size_t)i * np + j;
  if (tiled) {
    if (j >= np) return;
    cd v = {0, 0};
    if (i < n && j < n) v = cd{(double)negAB[e], (double)negAB[(size_t)np * np + e]};
    if (i == j) v.r += 1.0;
    S[e] = v.r;
    S[(size_t)np * np + e] = v.i;
    return;
  }
  if (j >= n) return;
  cd v = {(double)negAB[e], (double)negAB[(size_t)np * np + e]};
  if (i == j) v.r += 1.0;
  st_c(S, (size_t)i * n + j, v);
}

// out[q][i] = base[q][i] + sgn sum_j M[i][j] c[j] v[q][j] (i < n), one wave per
// (row, rhs). M is planar T with row stride ld (imaginary plane after ld^2
// entries) or, for T = double with ld == 0, row-major complex FP64 of order n.
// c (complex, optional) scales v; base (optional) is added.
template <typename T>
__global__ void __launch_bounds__(NT) k_wrap_gemv(int n, const T *__restrict__ Mt, int ld,
                                                  const double *__restrict__ c, const double *__restrict__ v,
                                                  size_t vstride, const double *__restrict__ base, size_t bstride,
                                                  double sgn, double *__restrict__ out, size_t ostride) {
  const int lane = threadIdx.x & 63, i = blockIdx.x * (NT / 64) + (threadIdx.x >> 6), q = blockIdx.y;
  if (i >= n) return;
  const double *vq = v + vstride * q;
  double sr = 0, si = 0;
  for (int j = lane; j < n; j += 64) {
    cd mv;
    if (ld > 0) {
      const size_t e = (size_t)i * ld + j;
      mv = cd{(double)Mt[e], (double)Mt[(size_t)ld * ld + e]};
    } else {
      mv = ld_c(reinterpret_cast<const double *>(Mt), (size_t)i * n + j);
    }
    cd x = ld_c(vq, j);
    if (c) x = cmul(ld_c(c, j), x);
    sr += mv.r * x.r - mv.i * x.i;
    si += mv.r * x.i + mv.i * x.r;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    sr += __shfl_xor(sr, off);
    si += __shfl_xor(si, off);
  }
  if (lane == 0) {
    cd r = {sgn * sr, sgn * si};
    if (base) r = cadd(r, ld_c(base + bstride * q, i));
    st_c(out + ostride * q, i, r);
  }
}

// Lagged truncated wrap (one sweep per refinement step). Per right-hand side
// q, five n-vectors sit in blocks 0..4 of the otherwise unused f buffer
// (plane q of [nrhs][m][n], m > 5 on this path): g0, gl, the boundary part of
// the residual left by the last step's coupling (blocks 0 and m-1); tl =
// G_{m-1} gl; f0, fl, the wrap forcing added to blocks 0 and m-1 of r.
enum { LAG_G0 = 0, LAG_GL = 1, LAG_TL = 2, LAG_F0 = 3, LAG_FL = 4 };

// f0 = -XM_0 w1, fl = -XP_{m-1} w2 (kept for the next step's g) and added to
// blocks 0 and m-1 of r
__global__ void __launch_bounds__(NT) k_wrap_f_lag(fdfd_grid g, const double *__restrict__ op,
                                                   const double *__restrict__ w1, const double *__restrict__ w2,
                                                   double *__restrict__ lag, double *__restrict__ r) {
  const int n = g.ny, m = g.nx, i = blockIdx.x * NT + threadIdx.x, q = blockIdx.y;
  if (i >= n) return;
  const size_t M = (size_t)m * n;
  const cd a = cmul(ld_c(op, FDFD_OP_XM * M + i), ld_c(w1 + 2 * (size_t)n * q, i));
  const cd b = cmul(ld_c(op, FDFD_OP_XP * M + (size_t)(m - 1) * n + i), ld_c(w2 + 2 * (size_t)n * q, i));
  double *lq = lag + 2 * M * q, *rq = r + 2 * M * q;
  st_c(lq, (size_t)LAG_F0 * n + i, cd{-a.r, -a.i});
  st_c(lq, (size_t)LAG_FL * n + i, cd{-b.r, -b.i});
  st_c(rq, i, csub(ld_c(rq, i), a));
  st_c(rq, (size_t)(m - 1) * n + i, csub(ld_c(rq, (size_t)(m - 1) * n + i), b));
}

// After the sweep s = T^-1 (r + f): the residual's boundary part is
// g0 = -(f0 + XM_0 s_{m-1}) on block 0 and gl = -(fl + XP_{m-1} s_0) on block
// m-1 (A s = T s + E s = r + f + E s). with_f = 0: no forcing was added.
__global__ void __launch_bounds__(NT) k_wrap_lag_next(fdfd_grid g, const double *__restrict__ op,
                                                      const double *__restrict__ s, int with_f,
                                                      double *__restrict__ lag) {
  const int n = g.ny, m = g.nx, i = blockIdx.x * NT + threadIdx.x, q = blockIdx.y;
  if (i >= n) return;
  const size_t M = (size_t)m * n;
  const double *sq = s + 2 * M * q;
  double *lq = lag + 2 * M * q;
  cd g0 = cmul(ld_c(op, FDFD_OP_XM * M + i), ld_c(sq, (size_t)(m - 1) * n + i));
  cd gl = cmul(ld_c(op, FDFD_OP_XP * M + (size_t)(m - 1) * n + i), ld_c(sq, i));
  if (with_f) {
    g0 = cadd(g0, ld_c(lq, (size_t)LAG_F0 * n + i));
    gl = cadd(gl, ld_c(lq, (size_t)LAG_FL * n + i));
  }
  st_c(lq, (size_t)LAG_G0 * n + i, cd{-g0.r, -g0.i});
  st_c(lq, (size_t)LAG_GL * n + i, cd{-gl.r, -gl.i});
}

// ------------------------------------------------------------------ norms
// part[r][blk] = sum |w v|^2 over a slice; w = stretch weight when weighted.
__device__ inline double stretch_abs2(const fdfd_grid &g, int i, int j) {
  const double hx = ((double)g.nx * g.dl) / g.nx, hy = ((double)g.ny * g.dl) / g.ny;
  cd s = cmul(sfactor(0, i, g.nx, g.npml_x, hx, g.omega, g.L0),
              sfactor(0, j, g.ny, g.npml_y, hy, g.omega, g.L0));
  return s.r * s.r + s.i * s.i;
}
__global__ void __launch_bounds__(NT) k_norm_part(fdfd_grid g, const double *__restrict__ v,
                                                  int weighted, double *__restrict__ part) {
  __shared__ double red[NT];
  const size_t M = (size_t)g.nx * g.ny;
  const int r = blockIdx.y;
  double acc = 0;
  for (size_t p = (size_t)blockIdx.x * NT + threadIdx.x; p < M; p += (size_t)gridDim.x * NT) {
    cd z = ld_c(v + 2 * M * r, p);
    double a2 = z.r * z.r + z.i * z.i;
    if (weighted) a2 *= stretch_abs2(g, (int)(p / g.ny), (int)(p % g.ny));
    acc += a2;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)r * gridDim.x + blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(NT) k_norm_final(const double *__restrict__ part, int nparts,
                                                   double *__restrict__ out) {
  __shared__ double red[NT];
  const int r = blockIdx.x;
  double acc = 0;
  for (int q = threadIdx.x; q < nparts; q += NT) acc += part[(size_t)r * nparts + q];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[r] = red[0];
}

// --------------------------------------------------- elementwise helpers
__global__ void __launch_bounds__(NT) k_stretch_scale(fdfd_grid g, double *v, int inverse) {
  const int j = blockIdx.x * NT + threadIdx.x, i = blockIdx.y, r = blockIdx.z;
  if (j >= g.ny) return;
  const double hx = ((double)g.nx * g.dl) / g.nx, hy = ((double)g.ny * g.dl) / g.ny;
  cd s = cmul(sfactor(0, i, g.nx, g.npml_x, hx, g.omega, g.L0),
              sfactor(0, j, g.ny, g.npml_y, hy, g.omega, g.L0));
  if (inverse) s = cinv(s);
  const size_t M = (size_t)g.nx * g.ny, p = (size_t)r * M + (size_t)i * g.ny + j;
  st_c(v, p, cmul(ld_c(v, p), s));
}

__global__ void __launch_bounds__(NT) k_grad_eps(fdfd_grid g, const double *__restrict__ lam,
                                                 const double *__restrict__ e, double *__restrict__ grad) {
  const size_t M = (size_t)g.nx * g.ny;
  const size_t p = (size_t)blockIdx.x * NT + threadIdx.x;
  if (p >= M) return;
  const double w2e = g.omega * g.omega * EPSILON_0 * g.L0;
  cd z = cmul(ld_c(lam, p), ld_c(e, p));
  grad[p] = -w2e * z.r;
}

__global__ void __launch_bounds__(NT) k_kerr(fdfd_grid g, double *__restrict__ op,
                                             const double *__restrict__ base_c,
                                             const double *__restrict__ e,
                                             const double *__restrict__ chi,
                                             double *__restrict__ eps_nl) {
  const size_t M = (size_t)g.nx * g.ny;
  const size_t p = (size_t)blockIdx.x * NT + threadIdx.x;
  if (p >= M) return;
  cd z = ld_c(e, p);
  const double enl = chi[p] * (z.r * z.r + z.i * z.i);
  cd c = ld_c(base_c, p);
  c.r += g.omega * g.omega * EPSILON_0 * g.L0 * enl;
  st_c(op, FDFD_OP_C * M + p, c);
  if (eps_nl) eps_nl[p] = enl;
}


// Secondary fields (simulation.py:133-175): Ez -> (Hx, Hy), Hz -> (Ex, Ey).
__global__ void __launch_bounds__(NT) k_fields(fdfd_grid g, const double *__restrict__ eps,
                                               const double *__restrict__ x, double *__restrict__ f1,
                                               double *__restrict__ f2, int averaging) {
  const int j = blockIdx.x * NT + threadIdx.x, i = blockIdx.y, r = blockIdx.z;
  if (j >= g.ny) return;
  const int nx = g.nx, ny = g.ny;
  const double hx = ((double)nx * g.dl) / nx, hy = ((double)ny * g.dl) / ny;
  const size_t M = (size_t)nx * ny, p = (size_t)i * ny + j;
  const double *xr = x + 2 * M * r;
  const int im = (i + nx - 1) % nx, jm = (j + ny - 1) % ny;
  cd X = ld_c(xr, p);
  // Dyb X and Dxb X (derivatives.py:7-34 with the 1/s stretch, pml.py:79-97)
  cd dyb = cmul(cinv(sfactor(1, j, ny, g.npml_y, hy, g.omega, g.L0)),
                cscale(csub(X, ld_c(xr, (size_t)i * ny + jm)), 1.0 / hy));
  cd dxb = cmul(cinv(sfactor(1, i, nx, g.npml_x, hx, g.omega, g.L0)),
                cscale(csub(X, ld_c(xr, (size_t)im * ny + j)), 1.0 / hx));
  const cd inv_iw = {0.0, -1.0 / g.omega};  // 1/(i omega)
  cd a, b;
  if (g.pol == FDFD_POL_EZ) {
    const cd c = cscale(inv_iw, 1.0 / (MU_0 * g.L0));
    a = cscale(cmul(c, dyb), -1.0);
    b = cmul(c, dxb);
  } else {
    const double eps0_ = EPSILON_0 * g.L0;
    cd ee = cscale(ld_c(eps, p), eps0_), ex = ee, ey = ee;
    if (averaging) {
      ex = cscale(cadd(cscale(ld_c(eps, (size_t)im * ny + j), eps0_), ee), 0.5);
      ey = cscale(cadd(cscale(ld_c(eps, (size_t)i * ny + jm), eps0_), ee), 0.5);
    }
    a = cmul(inv_iw, cmul(cinv(ey), dyb));
    b = cscale(cmul(inv_iw, cmul(cinv(ex), dxb)), -1.0);
  }
  st_c(f1 + 2 * M * r, p, a);
  st_c(f2 + 2 * M * r, p, b);
}

// ---------------------------------------------------------------- host side
int check_grid(const fdfd_grid *g) {
  if (!g) return fail(FDFD_ERR_ARG, "grid is NULL");
  if (g->nx < 2 || g->ny < 2) return fail(FDFD_ERR_ARG, "grid must be at least 2x2, got %dx%d", g->nx, g->ny);
  if (g->ny > MAX_BLOCK_N)
    return fail(FDFD_ERR_ARG, "block size ny=%d exceeds the supported %d (sweep along the shorter axis)", g->ny,
                MAX_BLOCK_N);
  if (g->npml_x < 0 || g->npml_y < 0) return fail(FDFD_ERR_ARG, "both elements of NPML must be >= 0");
  if (g->pol != FDFD_POL_EZ && g->pol != FDFD_POL_HZ) return fail(FDFD_ERR_ARG, "pol must be Ez or Hz");
  if (!(g->omega > 0) || !(g->dl > 0) || !(g->L0 > 0)) return fail(FDFD_ERR_ARG, "omega, dl and L0 must be positive");
  return FDFD_OK;
}
int check_prec(int precision) {
  const int base = precision & 0xff, flags = precision & ~0xff;
  if ((base != FDFD_PREC_FP64 && base != FDFD_PREC_C64) ||
      (flags & ~(FDFD_FACTOR_PIVOT | FDFD_FACTOR_STEPWISE)))
    return fail(FDFD_ERR_ARG, "precision must be FDFD_PREC_FP64 or FDFD_PREC_C64 (optionally | "
                "FDFD_FACTOR_PIVOT), got %d", precision);
  return FDFD_OK;
}
inline int base_prec(int precision) { return precision & 0xff; }
inline size_t elem_size(int precision) { return base_prec(precision) == FDFD_PREC_FP64 ? 8 : 4; }

struct FactorWork {
  void *W; void *Pinv; int *status;
  unsigned *rflag;                             // [2 chains][MAX_TILES] column flags of the step kernel
  void *xcol; void *pinv_all; unsigned *sync;  // persistent-kernel exchange (W doubles as xrow)
  double *kscratch;                            // ping-pong partner of K (order 2n, complex FP64)
};
// The workspace header: the status word, then the column flags of both
// chains; it is cleared as one block at the start of a factorization.
constexpr int MAX_TILES = MAX_BLOCK_N / TB;
constexpr size_t WORK_HEADER = 1024;
static_assert(512 + 2 * MAX_TILES * 4 <= WORK_HEADER, "column flags fit the header");
size_t persist_sync_words(int nt) { return 2 + (size_t)nt + 2 * (size_t)nt * nt; }
FactorWork factor_work_layout(const fdfd_grid *g, int precision, void *work) {
  const size_t np = pad_n(g->ny), nt = np / TB, es = elem_size(precision);
  char *base = (char *)work;
  FactorWork fw;
  fw.status = (int *)base;
  fw.rflag = (unsigned *)(base + 512);
  size_t off = WORK_HEADER;
  fw.Pinv = base + off;     off += align_up(2 * 2 * TB * TB * es, 256);
  fw.W = base + off;        off += align_up(2 * 2 * np * np * es, 256);  // also xrow (2 parities)
  fw.xcol = base + off;     off += align_up(2 * 2 * np * np * es, 256);
  fw.pinv_all = base + off; off += align_up(2 * nt * 2 * TB * TB * es, 256);
  fw.sync = (unsigned *)(base + off); off += align_up(4 * persist_sync_words((int)nt), 256);
  fw.kscratch = (double *)(base + off);
  return fw;
}
size_t factor_work_bytes(const fdfd_grid *g, int precision) {
  const size_t np = pad_n(g->ny), nt = np / TB, es = elem_size(precision), N = 2 * (size_t)g->ny;
  return WORK_HEADER + align_up(2 * 2 * TB * TB * es, 256) + 2 * align_up(2 * 2 * np * np * es, 256) +
         align_up(2 * nt * 2 * TB * TB * es, 256) + align_up(4 * persist_sync_words((int)nt), 256) +
         align_up(16 * N * N, 256);
}

struct SolveWork {
  double *r; void *y; double *part; double *norms; SweepSync *sync;
  unsigned tag_next;     // tag base of the next persistent sweep launch (host side)
  double *t, *f;         // T^-1 r; the exact wrap's -U C K^-1 V^T t, or the lagged wrap's vectors
  double *u, *w1, *w2;   // truncated wrap: [nrhs][n] complex each
};
SolveWork solve_work_layout(const fdfd_grid *g, int precision, int nrhs, void *work) {
  const size_t M = (size_t)g->nx * g->ny, np = pad_n(g->ny), es = elem_size(precision);
  char *base = (char *)work;
  SolveWork sw;
  sw.tag_next = 1;
  size_t off = 0;
  sw.r = (double *)(base + off); off += align_up(16 * M * nrhs, 256);
  sw.y = base + off;             off += align_up(2 * es * (size_t)g->nx * np * nrhs, 256);
  sw.part = (double *)(base + off); off += align_up(8 * (size_t)NORM_PARTS * nrhs * 2, 256);
  sw.norms = (double *)(base + off); off += align_up(8 * 4 * (size_t)nrhs, 256);
  sw.sync = (SweepSync *)(base + off); off += align_up(sizeof(SweepSync), 256);
  sw.t = (double *)(base + off);  off += align_up(16 * M * nrhs, 256);
  sw.f = (double *)(base + off);  off += align_up(16 * M * nrhs, 256);
  const size_t wb = align_up(16 * (size_t)g->ny * nrhs, 256);
  sw.u = (double *)(base + off);  off += wb;
  sw.w1 = (double *)(base + off); off += wb;
  sw.w2 = (double *)(base + off);
  return sw;
}
size_t solve_work_bytes(const fdfd_grid *g, int precision, int nrhs) {
  const size_t M = (size_t)g->nx * g->ny, np = pad_n(g->ny), es = elem_size(precision);
  return align_up(16 * M * nrhs, 256) + align_up(2 * es * (size_t)g->nx * np * nrhs, 256) +
         align_up(8 * (size_t)NORM_PARTS * nrhs * 2, 256) + align_up(8 * 4 * (size_t)nrhs, 256) +
         align_up(sizeof(SweepSync), 256) + 2 * align_up(16 * M * nrhs, 256) +
         3 * align_up(16 * (size_t)g->ny * nrhs, 256);
}

// Exact wrap correction (full chains, K of order 2n) or the truncated one;
// the factorization and the solve take the same decision.
inline bool wrap_exact(const fdfd_grid *g) {
  return g->npml_x < WRAP_MIN_PML || g->npml_y < WRAP_MIN_PML || g->npml_x + WRAP_EXTRA >= g->nx - 1;
}
// The twist block c of the factorization. Twisted (c < m - 1): a top chain
// eliminates blocks 0..c-1 downwards (G_k = S_k^-1), a bottom chain blocks
// m-1..c+1 upwards (G_k = S'_k^-1, S'_k = T_k - XP_k G_{k+1} XM_{k+1}), both
// in the same step launches, and block c takes both:
// G_c = (T_c - XM_c G_{c-1} XP_{c-1} - XP_c G_{c+1} XM_{c+1})^-1. The serial
// chain of a factorization and of a sweep is then about half as long. It
// needs the truncated wrap with room for both corner chains (L < c and
// L <= m - 2 - c); otherwise one chain over all blocks (c = m - 1).
inline int twist_block(const fdfd_grid *g) {
  const int m = g->nx;
  if (!g_twist.load() || wrap_exact(g)) return m - 1;
  const int c = (m + 1) / 2, L = g->npml_x + WRAP_EXTRA;
  return (L < c && L <= m - 2 - c) ? c : m - 1;
}
// The factor buffer: the block inverses G_k ([m][2][np][np] of T), then the
// wrap correction: K^-1 (row-major complex FP64, order 2n) when exact, else
// S^-1 (order n, planar and padded to np when tiled) followed by copies of
// Sigma_0 and of T^-1_{m-1,m-1} (planar T, np x np each).
size_t factor_blocks_bytes(const fdfd_grid *g, int precision) {
  const size_t np = pad_n(g->ny);
  return align_up((size_t)g->nx * 2 * np * np * elem_size(precision), 256);
}
size_t wrap_sinv_bytes(const fdfd_grid *g) {
  const size_t np = pad_n(g->ny);
  return align_up(16 * np * np, 256);
}
// S^-1 of the truncated wrap is formed by the tiled Gauss-Jordan steps (planar
// [2][np][np] complex FP64) when their two pivot-inverse slots fit the
// workspace's exchange area and S fits the K scratch, else by the
// column-by-column kernels (row-major).
inline bool wrap_sinv_tiled(const fdfd_grid *g, int precision) {
  const int np = pad_n(g->ny);
  return np <= 2 * g->ny && (precision == FDFD_PREC_FP64 || np / TB >= 2);
}
size_t factor_total_bytes(const fdfd_grid *g, int precision) {
  const size_t n = g->ny, np = pad_n(g->ny);
  return factor_blocks_bytes(g, precision) +
         std::max(64 * n * n, wrap_sinv_bytes(g) + 4 * np * np * elem_size(precision));
}
inline double *wrap_kinv(const fdfd_grid *g, int precision, const void *factor) {
  return (double *)((char *)factor + factor_blocks_bytes(g, precision));
}
inline void *wrap_sigma(const fdfd_grid *g, int precision, const void *factor) {
  return (char *)factor + factor_blocks_bytes(g, precision) + wrap_sinv_bytes(g);
}
// T^-1_{m-1,m-1} of the truncated wrap: G_{m-1} of a single chain, or the
// mirrored chain's Sigma'_{m-1} of a twisted factorization
inline void *wrap_sigma_last(const fdfd_grid *g, int precision, const void *factor) {
  const size_t np = pad_n(g->ny);
  return (char *)wrap_sigma(g, precision, factor) + 2 * np * np * elem_size(precision);
}

template <typename T, bool PIV>
int launch_step(const FBatch<T> &b, hipStream_t s) {
  FBatch<T> bl = b;
  const int grid = batch_layout(bl);
  hipLaunchKernelGGL((k_gj_step<T, PIV>), dim3(grid), dim3(NT), 0, s, bl);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

// Persistent factorization when the whole tile grid fits on the device at
// once; returns FDFD_ERR_HIP with *ran = false when it does not fit or when
// a wait timed out (the caller then uses the launch-per-step path).
template <typename T, bool PIV>
int factor_persistent(const fdfd_grid *g, const double *op, void *factor, void *work, hipStream_t s,
                      bool *ran) {
  *ran = false;
  const int n = g->ny, np = pad_n(n), nt = np / TB;
  int dev = 0, cus = 0, per_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_gj_persist<T, PIV>, NT, 0));
  // every workgroup must be resident at once; keep a quarter of the device
  // for other streams' work
  if ((long)nt * nt > (long)per_cu * cus * 3 / 4) return FDFD_OK;
  FactorWork fw = factor_work_layout(g, sizeof(T) == 8 ? FDFD_PREC_FP64 : FDFD_PREC_C64, work);
  HIP_TRY(hipMemsetAsync(fw.status, 0, sizeof(int), s));
  HIP_TRY(hipMemsetAsync(fw.sync, 0, 4 * persist_sync_words(nt), s));
  PJob<T> J;
  J.op = op; J.G = (T *)factor; J.xrow = (T *)fw.W; J.xcol = (T *)fw.xcol; J.pinv = (T *)fw.pinv_all;
  J.sync = fw.sync; J.status = fw.status;
  J.m = g->nx; J.n = n; J.np = np; J.nt = nt; J.cap = g_spin_cap.load();
  hipLaunchKernelGGL((k_gj_persist<T, PIV>), dim3(nt * nt), dim3(NT), 0, s, J);
  HIP_TRY(hipGetLastError());
  unsigned ab = 0;
  HIP_TRY(hipMemcpyAsync(&ab, fw.sync, sizeof ab, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *ran = ab == 0;
  return FDFD_OK;
}

// The chain jobs of one factorization: the whole chain, or the top and
// bottom chains of a twisted one. The bottom chain works in the second half
// of W (room for the persistent kernel's two exchange parities) and keeps
// its pivot inverses in the persistent kernel's pivot area, both unused by
// the launch-per-step path otherwise.
// J has room for two jobs; only J[0] is written for a single chain.
template <typename T>
int chain_jobs(const fdfd_grid *g, const double *op, void *factor, void *work, bool sym, FJob<T> *J) {
  const int n = g->ny, np = pad_n(n), nt = np / TB, m = g->nx, c = twist_block(g);
  FactorWork fw = factor_work_layout(g, sizeof(T) == 8 ? FDFD_PREC_FP64 : FDFD_PREC_C64, work);
  memset(J, 0, sizeof(FJob<T>));
  J[0].op = op; J[0].G = (T *)factor; J[0].W = (T *)fw.W; J[0].Pinv = (T *)fw.Pinv; J[0].status = fw.status;
  J[0].dsym = sym ? (T *)fw.xcol : nullptr;
  // the column buffers follow the symmetrizer in the persistent kernel's
  // column area (4 np + 2 x 2 np TB elements of its 4 np^2)
  J[0].rbuf = (T *)((char *)fw.xcol + align_up(4 * (size_t)np * sizeof(T), 256));
  J[0].rflag = (sym && nt > 1) ? fw.rflag : nullptr;
  J[0].m = m; J[0].n = n; J[0].np = np; J[0].nt = nt;
  J[0].k = -1; J[0].p = 0; J[0].dir = 1; J[0].c = c;
  if (c == m - 1) return 1;
  J[1] = J[0];
  J[1].W = (T *)fw.W + 2 * (size_t)np * np;
  J[1].Pinv = (T *)fw.pinv_all;
  J[1].rbuf = J[0].rbuf + 2 * (size_t)np * TB;
  J[1].rflag = J[0].rflag ? fw.rflag + MAX_TILES : nullptr;
  J[1].dir = -1;
  return 2;
}
// Steps of a chain job after its initial pivot: the top chain's
// (c + 1) nt; the bottom chain's (m - 1 - c) nt end earlier.
template <typename T>
inline int chain_steps(const FJob<T> &J) {
  return (J.dir > 0 ? J.c + 1 : J.m - 1 - J.c) * J.nt;
}
template <typename T>
inline void chain_at(FJob<T> &J, int st) {
  J.p = st % J.nt;
  J.k = J.dir > 0 ? st / J.nt : J.m - 1 - st / J.nt;
}

template <typename T, bool PIV>
int factor_impl(const fdfd_grid *g, const double *op, void *factor, void *work, hipStream_t s) {
  FactorWork fw = factor_work_layout(g, sizeof(T) == 8 ? FDFD_PREC_FP64 : FDFD_PREC_C64, work);
  HIP_TRY(hipMemsetAsync(fw.status, 0, WORK_HEADER, s));
  FBatch<T> b;
  memset(&b, 0, sizeof b);
  const int nch = chain_jobs<T>(g, op, factor, work, g_step_sym.load() != 0, b.job);
  b.njobs = nch;
  int rc = launch_step<T, PIV>(b, s);
  if (rc) return rc;
  const int steps = chain_steps(b.job[0]);
  for (int st = 0; st < steps; ++st) {
    b.njobs = 0;
    for (int i = 0; i < nch; ++i)
      if (st < chain_steps(b.job[i])) {
        chain_at(b.job[i], st);
        b.njobs = i + 1;
      }
    if ((rc = launch_step<T, PIV>(b, s))) return rc;
  }
  return FDFD_OK;
}

// The block factorizations of njobs operators of one grid in shared step
// launches (the request loop's batched launches, on the caller's stream).
template <typename T, bool PIV>
int factor_batch_impl(const fdfd_grid *g, const double *const *ops, void *const *factors, void *const *works,
                      int njobs, hipStream_t s) {
  FBatch<T> b;
  memset(&b, 0, sizeof b);
  FJob<T> top[MAX_BATCH], bot[MAX_BATCH];
  int nch = 1;
  for (int j = 0; j < njobs; ++j) {
    FactorWork fw = factor_work_layout(g, sizeof(T) == 8 ? FDFD_PREC_FP64 : FDFD_PREC_C64, works[j]);
    HIP_TRY(hipMemsetAsync(fw.status, 0, WORK_HEADER, s));
    FJob<T> pair[2];
    nch = chain_jobs<T>(g, ops[j], factors[j], works[j], g_step_sym.load() != 0, pair);
    top[j] = pair[0];
    bot[j] = pair[1];
  }
  // every chain's initial pivot, then the steps: the top chains (and, while
  // they last, the bottom chains) of all operators in each launch
  auto fill = [&](int st) {
    b.njobs = 0;
    for (int j = 0; j < njobs; ++j) {
      if (st >= 0) chain_at(top[j], st);
      b.job[b.njobs++] = top[j];
    }
    if (nch == 2 && (st < 0 || st < chain_steps(bot[0])))
      for (int j = 0; j < njobs; ++j) {
        if (st >= 0) chain_at(bot[j], st);
        b.job[b.njobs++] = bot[j];
      }
  };
  fill(-1);
  int rc = launch_step<T, PIV>(b, s);
  if (rc) return rc;
  const int steps = chain_steps(top[0]);
  for (int st = 0; st < steps; ++st) {
    fill(st);
    if ((rc = launch_step<T, PIV>(b, s))) return rc;
  }
  return FDFD_OK;
}

template <typename T>
void sweep_block_attr() {
  static std::once_flag lds_attr;
  std::call_once(lds_attr, [] {
    const int most = (int)(2 * MAXR * pad_n(MAX_BLOCK_N) * sizeof(T));
    (void)hipFuncSetAttribute((const void *)k_sweep_fwd<T>, hipFuncAttributeMaxDynamicSharedMemorySize, most);
    (void)hipFuncSetAttribute((const void *)k_sweep_bwd<T>, hipFuncAttributeMaxDynamicSharedMemorySize, most);
  });
}

// y_c and x_c += y_c of the twist block, between the forward chains and the
// backward chains of a twisted sweep
template <typename T>
void launch_twist_block(const fdfd_grid *g, const T *G, const double *op, const double *r, T *y, double *x,
                        int nrhs, int c, hipStream_t s) {
  const int np = pad_n(g->ny);
  sweep_block_attr<T>();
  hipLaunchKernelGGL(k_sweep_fwd<T>, dim3((np + 3) / 4), dim3(NT), (size_t)2 * MAXR * np * sizeof(T), s, G, op, r,
                     y, x, g->nx, g->ny, np, c, nrhs, c - 1, (int)FDFD_OP_XM, c + 1, (int)FDFD_OP_XP);
}

template <typename T>
int sweep_blockwise(const fdfd_grid *g, const double *op, const void *factor, const double *r,
                    double *x, int nrhs, void *y, hipStream_t s) {
  const int nx = g->nx, ny = g->ny, np = pad_n(ny), c = twist_block(g);
  const dim3 grid((np + 3) / 4);
  const size_t lds = (size_t)2 * MAXR * np * sizeof(T);
  const T *G = (const T *)factor;
  T *yt = (T *)y;
  sweep_block_attr<T>();
  auto fwd = [&](int k, int nb, int pl) {
    hipLaunchKernelGGL(k_sweep_fwd<T>, grid, dim3(NT), lds, s, G, op, r, yt, (double *)nullptr, nx, ny, np, k,
                       nrhs, nb, pl, -1, 0);
  };
  auto bwd = [&](int k, int nb, int pl) {
    hipLaunchKernelGGL(k_sweep_bwd<T>, grid, dim3(NT), lds, s, G, op, yt, x, nx, ny, np, k, nrhs, nb, pl);
  };
  if (c == nx - 1) {
    for (int k = 0; k < nx; ++k) fwd(k, k > 0 ? k - 1 : -1, FDFD_OP_XM);
    for (int k = nx - 1; k >= 0; --k) bwd(k, k + 1 < nx ? k + 1 : -1, FDFD_OP_XP);
  } else {
    for (int k = 0; k < c; ++k) fwd(k, k > 0 ? k - 1 : -1, FDFD_OP_XM);
    for (int k = nx - 1; k > c; --k) fwd(k, k + 1 < nx ? k + 1 : -1, FDFD_OP_XP);
    launch_twist_block<T>(g, G, op, r, yt, x, nrhs, c, s);
    for (int k = c - 1; k >= 0; --k) bwd(k, k + 1, FDFD_OP_XP);
    for (int k = c + 1; k < nx; ++k) bwd(k, k - 1, FDFD_OP_XM);
  }
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

// The wrap correction of a finished block factorization: the two backward
// chains (two launches per block), K, and its inverse (2 launches per column
// of K). Temporaries live in the factorization workspace, whose exchange
// buffers are free once the block inverses are complete.
template <typename T>
int wrap_setup(const fdfd_grid *g, const double *op, void *factor, void *work, bool pivoted, hipStream_t s) {
  const int prec = sizeof(T) == 8 ? FDFD_PREC_FP64 : FDFD_PREC_C64;
  const int m = g->nx, n = g->ny, np = pad_n(n), nt = np / TB;
  const size_t np2 = (size_t)np * np, M = (size_t)m * n;
  FactorWork fw = factor_work_layout(g, prec, work);
  const T *G = (const T *)factor;
  T *Sig = (T *)fw.W, *Q = (T *)fw.W + 2 * np2;
  T *Z[2] = {(T *)fw.xcol, (T *)fw.xcol + 2 * np2};
  const T *Glast = G + (size_t)(m - 1) * 2 * np2;
  // exact chains over all blocks, or the truncated Sigma chain without Z;
  // for a twisted factorization also the mirrored truncated chain over the
  // bottom blocks, Sigma'_{m-1-L} = G_{m-1-L},
  // Sigma'_k = G_k + G_k XM_k Sigma'_{k-1} XP_{k-1} G_k, in the same launches
  // (its temporaries in the second exchange buffer, free after the
  // factorization)
  const bool exact = wrap_exact(g);
  const bool twisted = twist_block(g) < m - 1;
  const int L = exact ? m - 1 : g->npml_x + WRAP_EXTRA;
  const T *sig_prev = G + (size_t)L * 2 * np2, *z_prev = exact ? Glast : nullptr;
  T *SigB = (T *)fw.xcol, *QB = (T *)fw.xcol + 2 * np2;
  const T *sigb_prev = G + (size_t)(m - 1 - L) * 2 * np2;
  // Sigma_k and Sigma'_k are D-symmetric like G_k: the products that form
  // them compute their upper tiles only. D goes to the pivot area of the
  // workspace (free once the block factorization is done).
  T *dsym = (T *)fw.pinv_all;
  hipLaunchKernelGGL(k_block_symmetrizer<T>, dim3(1), dim3(64), 0, s, op, m, n, np, dsym);
  int zi = 0;
  for (int i = 0; i < L; ++i) {
    const int k = L - 1 - i, kb = m - L + i;   // top and bottom chain blocks
    const T *Gk = G + (size_t)k * 2 * np2, *Gb = G + (size_t)kb * 2 * np2;
    GemmBatch<T> b1;
    memset(&b1, 0, sizeof b1);
    b1.njobs = 1;
    // Q = G_k (XP_k Sigma_{k+1} XM_{k+1})
    b1.job[0] = GemmJob<T>{Gk, sig_prev, nullptr, Q, op + 2 * (FDFD_OP_XP * M + (size_t)k * n),
                           op + 2 * (FDFD_OP_XM * M + (size_t)(k + 1) * n), T(1), nullptr};
    // Z_k = -G_k (XP_k Z_{k+1})
    if (exact)
      b1.job[b1.njobs++] = GemmJob<T>{Gk, z_prev, nullptr, Z[zi], op + 2 * (FDFD_OP_XP * M + (size_t)k * n),
                                      nullptr, T(-1), nullptr};
    // Q' = G_kb (XM_kb Sigma'_{kb-1} XP_{kb-1})
    if (twisted)
      b1.job[b1.njobs++] = GemmJob<T>{Gb, sigb_prev, nullptr, QB, op + 2 * (FDFD_OP_XM * M + (size_t)kb * n),
                                      op + 2 * (FDFD_OP_XP * M + (size_t)(kb - 1) * n), T(1), nullptr};
    hipLaunchKernelGGL(k_cgemm<T>, dim3(b1.njobs * nt * nt), dim3(GT), 0, s, b1, n, np);
    // Sigma_k = G_k + Q G_k (and Sigma'_kb = G_kb + Q' G_kb)
    GemmBatch<T> b2;
    memset(&b2, 0, sizeof b2);
    b2.njobs = 1;
    b2.job[0] = GemmJob<T>{Q, Gk, Gk, Sig, nullptr, nullptr, T(1), dsym};
    if (twisted) b2.job[b2.njobs++] = GemmJob<T>{QB, Gb, Gb, SigB, nullptr, nullptr, T(1), dsym};
    hipLaunchKernelGGL(k_cgemm<T>, dim3(b2.njobs * nt * (nt + 1) / 2), dim3(GT), 0, s, b2, n, np);
    sig_prev = Sig;
    sigb_prev = SigB;
    if (exact) {
      z_prev = Z[zi];
      zi ^= 1;
    }
  }
  HIP_TRY(hipGetLastError());
  double *K = wrap_kinv(g, prec, factor);
  const int N = exact ? 2 * n : n;   // order of the matrix inverted below
  if (exact) {
    hipLaunchKernelGGL(k_wrap_capacitance<T>, dim3((N + NT - 1) / NT, N), dim3(NT), 0, s, *g, op, Glast, sig_prev,
                       z_prev, np, K);
  } else {
    // copies of Sigma_0 and of T^-1_{m-1,m-1} (G_{m-1}, or Sigma'_{m-1} of a
    // twisted factorization)
    T *sig0 = (T *)wrap_sigma(g, prec, factor), *sigl = (T *)wrap_sigma_last(g, prec, factor);
    HIP_TRY(hipMemcpyAsync(sig0, sig_prev, 2 * np2 * sizeof(T), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(sigl, twisted ? sigb_prev : Glast, 2 * np2 * sizeof(T), hipMemcpyDeviceToDevice, s));
    // S = I - Sigma_0 XM_0 T^-1_{m-1,m-1} XP_{m-1} (order n)
    GemmBatch<T> b3;
    memset(&b3, 0, sizeof b3);
    b3.njobs = 1;
    b3.job[0] = GemmJob<T>{sig0, sigl, nullptr, Q, op + 2 * (FDFD_OP_XM * M),
                           op + 2 * (FDFD_OP_XP * M + (size_t)(m - 1) * n), T(-1), nullptr};
    hipLaunchKernelGGL(k_cgemm<T>, dim3(nt * nt), dim3(GT), 0, s, b3, n, np);
    const bool tiled = wrap_sinv_tiled(g, prec);
    if (tiled) {
      // S into the K scratch; its inverse into the S^-1 slot of the factor
      hipLaunchKernelGGL(k_wrap_schur<T>, dim3((np + NT - 1) / NT, np), dim3(NT), 0, s, (const T *)Q, np, n,
                         fw.kscratch, 1);
    } else {
      HIP_TRY(hipMemsetAsync(K, 0, wrap_sinv_bytes(g), s));
      hipLaunchKernelGGL(k_wrap_schur<T>, dim3((n + NT - 1) / NT, n), dim3(NT), 0, s, (const T *)Q, np, n, K, 0);
    }
    // the rest of the tail (sized for the exact form) is cleared, so a factor
    // buffer is a deterministic function of the operator
    const size_t used = factor_blocks_bytes(g, prec) + wrap_sinv_bytes(g) + 4 * np2 * sizeof(T);
    const size_t total = factor_total_bytes(g, prec);
    if (total > used) HIP_TRY(hipMemsetAsync((char *)factor + used, 0, total - used, s));
    if (tiled) {
      // np/32 + 1 step launches of the block factorization's kernel on one
      // dense FP64 job: diagonal 32 x 32 pivot tiles (S is a small
      // perturbation of I), partial pivoting inside them for a pivoted factor
      FBatch<double> b;
      memset(&b, 0, sizeof b);
      b.njobs = 1;
      FJob<double> &J = b.job[0];
      J.op = op; J.G = K; J.W = (double *)fw.W; J.Pinv = (double *)fw.pinv_all; J.status = fw.status;
      J.dense = fw.kscratch; J.dsym = nullptr;
      J.m = 1; J.n = n; J.np = np; J.nt = nt;
      J.k = -1; J.p = 0; J.dir = 1; J.c = 0;
      int rc = pivoted ? launch_step<double, true>(b, s) : launch_step<double, false>(b, s);
      for (int p = 0; p < nt && !rc; ++p) {
        J.k = 0; J.p = p;
        rc = pivoted ? launch_step<double, true>(b, s) : launch_step<double, false>(b, s);
      }
      return rc;
    }
  }
  // Gauss-Jordan: two contiguous pivot-column copies and the pivot record in
  // the (now free) pivot-inverse area, the ping-pong partner of K in the
  // workspace's K scratch
  double *colb[2] = {(double *)fw.pinv_all, (double *)fw.pinv_all + 2 * (size_t)N};
  int *piv = (int *)((double *)fw.pinv_all + 4 * (size_t)N);
  double *Kb[2] = {K, fw.kscratch};
  const dim3 kgrid((N + NT - 1) / NT, (N + KGJ_ROWS - 1) / KGJ_ROWS);
  hipLaunchKernelGGL(k_kgj_col0, dim3((N + NT - 1) / NT), dim3(NT), 0, s, (const double *)K, N, colb[0]);
  for (int j = 0; j < N; ++j)
    hipLaunchKernelGGL(k_kgj_col, kgrid, dim3(NT), 0, s, (const double *)Kb[j & 1], Kb[(j + 1) & 1], N, j,
                       (const double *)colb[j & 1], colb[(j + 1) & 1], piv, fw.status);
  if (N & 1) HIP_TRY(hipMemcpyAsync(K, fw.kscratch, 16 * (size_t)N * N, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_kgj_unpermute, dim3((N + NT - 1) / NT), dim3(NT), 0, s, K, N, (const int *)piv);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

// The truncated wrap correction, lagged: the forcing f that makes
// T^-1 (r + f) correct the boundary part g of r. For t = T^-1 g,
// f = -U C K^-1 V^T t through S^-1: u = t_0 - Sigma_0 (XM_0 t_{m-1}),
// w2 = S^-1 u, w1 = t_{m-1} - G_{m-1} (XP_{m-1} w2), f = (-XM_0 w1 on block 0,
// -XP_{m-1} w2 on block m-1), with t's boundary blocks taken as Sigma_0 g0
// and G_{m-1} gl (the far corners dropped); f goes into r.
template <typename T>
int wrap_lag_forcing(const fdfd_grid *g, const double *op, const void *factor, const SolveWork &sw, int nrhs,
                     double *r, hipStream_t s) {
  const int prec = sizeof(T) == 8 ? FDFD_PREC_FP64 : FDFD_PREC_C64;
  const int m = g->nx, n = g->ny, np = pad_n(n);
  const size_t M = (size_t)m * n, nv = 2 * (size_t)n;
  const T *sigma = (const T *)wrap_sigma(g, prec, factor);
  const T *Glast = (const T *)wrap_sigma_last(g, prec, factor);   // T^-1_{m-1,m-1}
  const double *sinv = wrap_kinv(g, prec, factor);
  const double *XM0 = op + 2 * (FDFD_OP_XM * M), *XPl = op + 2 * (FDFD_OP_XP * M + (size_t)(m - 1) * n);
  double *lag = sw.f;
  const double *g0 = lag + LAG_G0 * nv, *gl = lag + LAG_GL * nv;
  double *tl = lag + LAG_TL * nv;
  const dim3 gg((n + NT / 64 - 1) / (NT / 64), nrhs);
  // tl = G_{m-1} gl
  hipLaunchKernelGGL(k_wrap_gemv<T>, gg, dim3(NT), 0, s, n, Glast, np, (const double *)nullptr, gl, 2 * M,
                     (const double *)nullptr, (size_t)0, 1.0, tl, 2 * M);
  // u = Sigma_0 g0 - Sigma_0 (XM_0 tl)   (w1 holds Sigma_0 g0)
  hipLaunchKernelGGL(k_wrap_gemv<T>, gg, dim3(NT), 0, s, n, sigma, np, (const double *)nullptr, g0, 2 * M,
                     (const double *)nullptr, (size_t)0, 1.0, sw.w1, nv);
  hipLaunchKernelGGL(k_wrap_gemv<T>, gg, dim3(NT), 0, s, n, sigma, np, XM0, (const double *)tl, 2 * M,
                     (const double *)sw.w1, nv, -1.0, sw.u, nv);
  // w2 = S^-1 u; w1 = tl - G_{m-1} (XP_{m-1} w2)
  hipLaunchKernelGGL(k_wrap_gemv<double>, gg, dim3(NT), 0, s, n, sinv, wrap_sinv_tiled(g, prec) ? np : 0,
                     (const double *)nullptr, (const double *)sw.u, nv, (const double *)nullptr, (size_t)0, 1.0,
                     sw.w2, nv);
  hipLaunchKernelGGL(k_wrap_gemv<T>, gg, dim3(NT), 0, s, n, Glast, np, XPl, (const double *)sw.w2, nv,
                     (const double *)tl, 2 * M, -1.0, sw.w1, nv);
  hipLaunchKernelGGL(k_wrap_f_lag, dim3((n + NT - 1) / NT, nrhs), dim3(NT), 0, s, *g, op, (const double *)sw.w1,
                     (const double *)sw.w2, lag, r);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

template <typename T, int RP>
void launch_psweep(int W, size_t lds, hipStream_t s, const T *G, const double *op, const double *r, T *y,
                   double *x, int nx, int ny, int np, int nr, int dir, const SweepChains &chs, SweepSync *sy,
                   unsigned tag0) {
  static std::once_flag lds_attr;
  std::call_once(lds_attr, [] {
    if constexpr (RP <= LL_MAXRHS)
      (void)hipFuncSetAttribute((const void *)k_psweep_ll<T, RP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(2 * RP * PSWEEP_MAX_N * sizeof(T)));
    else
      (void)hipFuncSetAttribute((const void *)k_psweep<T, RP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(2 * RP * PSWEEP_MAX_N * sizeof(T)));
  });
  if constexpr (RP <= LL_MAXRHS)
    hipLaunchKernelGGL((k_psweep_ll<T, RP>), dim3(W * chs.nch), dim3(NT), lds, s, G, op, r, y, x, nx, ny, np, nr,
                       dir, chs, sy, g_spin_cap.load(), tag0);
  else
    hipLaunchKernelGGL((k_psweep<T, RP>), dim3(W * chs.nch), dim3(NT), lds, s, G, op, r, y, x, nx, ny, np, nr,
                       dir, chs, sy, g_spin_cap.load());
}

// The chains of a persistent sweep: dir 0 forward, 1 backward.
SweepChains sweep_chains(const fdfd_grid *g, int dir) {
  const int m = g->nx, c = twist_block(g);
  SweepChains chs;
  memset(&chs, 0, sizeof chs);
  if (c == m - 1) {
    chs.nch = 1;
    chs.ch[0] = dir == 0 ? SweepChain{0, 1, m, FDFD_OP_XM, 0, 0} : SweepChain{m - 1, -1, m, FDFD_OP_XP, 0, 0};
  } else if (dir == 0) {
    chs.nch = 2;
    chs.ch[0] = SweepChain{0, 1, c, FDFD_OP_XM, 0, 0};
    chs.ch[1] = SweepChain{m - 1, -1, m - 1 - c, FDFD_OP_XP, 0, 0};
  } else {
    chs.nch = 2;
    chs.ch[0] = SweepChain{c - 1, -1, c, FDFD_OP_XP, 1, c};
    chs.ch[1] = SweepChain{c + 1, 1, m - 1 - c, FDFD_OP_XM, 1, c};
  }
  return chs;
}

// x += F^-1 r with the persistent sweep kernels (RHS in passes of
// sweep_rhs<T>()); blockwise = true uses one launch per block instead.
template <typename T>
int sweep_impl(const fdfd_grid *g, const double *op, const void *factor, const double *r,
               double *x, int nrhs, SolveWork &sw, hipStream_t s, bool blockwise) {
  // blocks beyond the persistent kernel's register tiles: one launch per block
  if (blockwise || pad_n(g->ny) > PSWEEP_MAX_N) return sweep_blockwise<T>(g, op, factor, r, x, nrhs, sw.y, s);
  const int nx = g->nx, ny = g->ny, np = pad_n(ny), c = twist_block(g);
  const int W = (np + 4 * SRW - 1) / (4 * SRW);
  const size_t M = (size_t)nx * ny;
  const SweepChains chs[2] = {sweep_chains(g, 0), sweep_chains(g, 1)};
  for (int q0 = 0; q0 < nrhs; q0 += sweep_rhs<T>()) {
    const int nr = std::min(sweep_rhs<T>(), nrhs - q0);
    const int rp = nr <= 1 ? 1 : nr <= 2 ? 2 : nr <= 4 ? 4 : 8;
    T *yq = (T *)sw.y + (size_t)2 * q0 * nx * np;
    const size_t lds = (size_t)2 * rp * pad64(np) * sizeof(T);
    // the abort flag is cleared by the caller, so an abort stays visible
    // across the sweeps of one solve attempt. Flag sweep: the step flags
    // start at zero. Tagged-word sweep: each launch's tags start past every
    // tag written before in this workspace since it was cleared.
    if (rp > LL_MAXRHS)
      for (int i = 0; i < chs[0].nch; ++i) HIP_TRY(hipMemsetAsync(sw.sync->ch[i].done, 0, sizeof sw.sync->ch[i].done, s));
    for (int dir = 0; dir < 2; ++dir) {
      if (sw.tag_next > 0xffffffffu - (unsigned)nx - 2) {
        for (int i = 0; i < 2; ++i) HIP_TRY(hipMemsetAsync(sw.sync->ch[i].llbuf, 0, sizeof sw.sync->ch[i].llbuf, s));
        sw.tag_next = 1;
      }
      const unsigned tag0 = sw.tag_next;
      sw.tag_next += (unsigned)nx + 1;
      const double *rq = dir == 0 ? r + 2 * M * q0 : nullptr;
      double *xq = dir == 1 ? x + 2 * M * q0 : nullptr;
      const T *G = (const T *)factor;
      const SweepChains &ch = chs[dir];
      switch (rp) {
        case 1: launch_psweep<T, 1>(W, lds, s, G, op, rq, yq, xq, nx, ny, np, nr, dir, ch, sw.sync, tag0); break;
        case 2: launch_psweep<T, 2>(W, lds, s, G, op, rq, yq, xq, nx, ny, np, nr, dir, ch, sw.sync, tag0); break;
        case 4: launch_psweep<T, 4>(W, lds, s, G, op, rq, yq, xq, nx, ny, np, nr, dir, ch, sw.sync, tag0); break;
        default:
          if constexpr (sweep_rhs<T>() >= 8)
            launch_psweep<T, 8>(W, lds, s, G, op, rq, yq, xq, nx, ny, np, nr, dir, ch, sw.sync, tag0);
          break;
      }
      // twisted: the twist block between the two passes
      if (dir == 0 && c < nx - 1)
        launch_twist_block<T>(g, G, op, r + 2 * M * q0, yq, x + 2 * M * q0, nr, c, s);
    }
  }
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

int norms_async(const fdfd_grid *g, const double *v, int nrhs, int weighted, double *part,
                double *out, hipStream_t s) {
  const size_t M = (size_t)g->nx * g->ny;
  int parts = (int)std::min<size_t>(NORM_PARTS, (M + NT - 1) / NT);
  hipLaunchKernelGGL(k_norm_part, dim3(parts, nrhs), dim3(NT), 0, s, *g, v, weighted, part);
  hipLaunchKernelGGL(k_norm_final, dim3(nrhs), dim3(NT), 0, s, (const double *)part, parts, out);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

// ------------------------------------------------------- batching engine
// A scheduler thread owns one stream and advances every submitted
// factorization by one Gauss-Jordan step per launch: the jobs in flight
// share each k_gj_step launch (up to MAXJ per launch, one precision per
// launch), new jobs join at the next launch and finished jobs leave, so
// factorizations of different sizes progress together.
struct EngineJob {
  fdfd_grid g;
  const double *op;
  void *factor;
  void *work;
  int precision;               // FDFD_PREC_*
  bool pivot;                  // partial pivoting inside pivot tiles
  bool sym = true;             // symmetric steps (upper tiles only)
  int m, n, np, nt;
  int step = -1;              // next step of the chains; -1: the initial pivots
  int nch = 1, steps = 0;     // chains (2: twisted) and the top chain's steps
  bool launched_all = false;  // every step is on the stream
  hipEvent_t done = nullptr;
  int rc = FDFD_OK;
};

}  // namespace

struct fdfd_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<EngineJob>> pending;
  std::vector<std::shared_ptr<EngineJob>> active;
  std::map<uint64_t, std::shared_ptr<EngineJob>> jobs;
  uint64_t next_ticket = 1;
  bool stop = false;
  int last_prec = 0;
  uint64_t launches = 0;
};

namespace {

// The chain jobs of an engine job at its next step into J (room for e.nch
// jobs); returns how many take part.
template <typename T>
int fill_jobs(FJob<T> *J, const EngineJob &e) {
  FJob<T> pair[2];
  chain_jobs<T>(&e.g, e.op, e.factor, e.work, e.sym, pair);
  int n = 0;
  for (int i = 0; i < e.nch; ++i) {
    if (e.step >= 0) {
      if (e.step >= chain_steps(pair[i])) break;   // the bottom chain ends first
      chain_at(pair[i], e.step);
    }
    J[n++] = pair[i];
  }
  return n;
}

// One batched launch over up to MAXJ jobs of precision T; advances their
// steps. Only the scheduler thread touches a job's k and p. A job whose last
// step is on the stream gets its completion event recorded first; then,
// under E->mu, its rc and launched_all are set (waiters read them under the
// same lock), so no waiter can see launched_all before the event exists on
// the stream.
template <typename T, bool PIV>
int engine_launch(fdfd_engine *E, std::vector<std::shared_ptr<EngineJob>> &sel) {
  FBatch<T> b;
  memset(&b, 0, sizeof b);
  for (auto &e : sel) b.njobs += fill_jobs<T>(b.job + b.njobs, *e);
  const int wg = batch_layout(b);
  hipLaunchKernelGGL((k_gj_step<T, PIV>), dim3(wg), dim3(NT), 0, E->stream, b);
  const hipError_t err = hipGetLastError();
  std::vector<std::pair<EngineJob *, int>> finished;  // job, rc
  for (auto &e : sel) {
    if (err != hipSuccess) { finished.emplace_back(e.get(), FDFD_ERR_HIP); continue; }
    if (++e->step == e->steps)
      finished.emplace_back(e.get(), hipEventRecord(e->done, E->stream) == hipSuccess ? FDFD_OK : FDFD_ERR_HIP);
  }
  std::lock_guard<std::mutex> lk(E->mu);
  for (auto &f : finished) {
    f.first->rc = f.second;
    f.first->launched_all = true;
  }
  ++E->launches;
  return err == hipSuccess ? FDFD_OK : FDFD_ERR_HIP;
}

void engine_loop(fdfd_engine *E) {
  (void)hipSetDevice(E->device);
  for (;;) {
    std::vector<std::shared_ptr<EngineJob>> sel;
    int prec = 0;
    {
      std::unique_lock<std::mutex> lk(E->mu);
      E->cv_work.wait(lk, [E] { return E->stop || !E->pending.empty() || !E->active.empty(); });
      while (!E->pending.empty()) {
        auto e = E->pending.front();
        E->pending.pop_front();
        FactorWork fw = factor_work_layout(&e->g, e->precision, e->work);
        if (hipMemsetAsync(fw.status, 0, WORK_HEADER, E->stream) != hipSuccess) e->rc = FDFD_ERR_HIP;
        E->active.push_back(e);
      }
      if (E->active.empty()) {
        if (E->stop) return;
        continue;
      }
      // one launch per kind (precision x pivoting); kinds take turns
      bool has[4] = {false, false, false, false};
      for (auto &e : E->active) has[e->precision * 2 + (e->pivot ? 1 : 0)] = true;
      prec = E->last_prec;
      for (int t = 1; t <= 4; ++t)
        if (has[(E->last_prec + t) % 4]) { prec = (E->last_prec + t) % 4; break; }
      E->last_prec = prec;
      int slots = 0;   // chains in the launch
      for (auto &e : E->active)
        if (e->precision * 2 + (e->pivot ? 1 : 0) == prec && slots + e->nch <= MAXJ) {
          sel.push_back(e);
          slots += e->nch;
        }
    }
    switch (prec) {
      case 0: engine_launch<double, false>(E, sel); break;
      case 1: engine_launch<double, true>(E, sel); break;
      case 2: engine_launch<float, false>(E, sel); break;
      default: engine_launch<float, true>(E, sel); break;
    }
    {
      std::lock_guard<std::mutex> lk(E->mu);
      bool any = false;
      for (auto it = E->active.begin(); it != E->active.end();) {
        if ((*it)->launched_all) { it = E->active.erase(it); any = true; }
        else ++it;
      }
      // round robin: jobs that just ran go to the back
      if (!any && E->active.size() > sel.size())
        std::rotate(E->active.begin(), E->active.begin() + std::min(sel.size(), E->active.size()),
                    E->active.end());
      if (any) E->cv_done.notify_all();
    }
  }
}

}  // namespace

// ------------------------------------------ flux probe (simulation.py:260-315)
// out = dl * sum over the line of the time-averaged Poynting flux. kind 0:
// Ez, x-normal (-Re(avg_x Ez conj Hy) / 2); 1: Ez, y-normal (Re(avg_y Ez
// conj Hx) / 2); 2: Hz, x-normal (Re(Ey conj avg_x Hz) / 2); 3: Hz, y-normal
// (-Re(Ex conj avg_y Hz) / 2). avg_x f at (i, j) is (f(i, j) + f(i+1, j)) / 2,
// avg_y likewise along j. One workgroup, fixed-order reduction.
__global__ void __launch_bounds__(NT) k_flux(int ny, double dl, int kind, const double *__restrict__ e,
                                             const double *__restrict__ h, int i0, int j0, int len,
                                             double *__restrict__ out) {
  __shared__ double part[NT];
  const bool xn = kind == 0 || kind == 2;
  const double *avg_src = kind < 2 ? e : h;   // the field averaged across the line
  const double *other = kind < 2 ? h : e;
  double acc = 0;
  for (int t = threadIdx.x; t < len; t += NT) {
    const int i = xn ? i0 : i0 + t, j = xn ? j0 + t : j0;
    const size_t p = (size_t)i * ny + j, q = xn ? p + ny : p + 1;
    const cd a = cscale(cadd(ld_c(avg_src, p), ld_c(avg_src, q)), 0.5);
    const cd o = ld_c(other, p);
    // Re(u conj v) = ur vr + ui vi; the field order only changes the sign
    const double re = a.r * o.r + a.i * o.i;
    acc += (kind == 0 || kind == 3) ? -0.5 * re : 0.5 * re;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = dl * part[0];
}

// =================================================================== C ABI
extern "C" {

int fdfd_abi_version(void) { return FDFD_ABI_VERSION; }
const char *fdfd_last_error(void) { return g_err.c_str(); }
int fdfd_block_padding(void) { return TB; }

size_t fdfd_op_bytes(const fdfd_grid *g) {
  if (check_grid(g)) return 0;
  return (size_t)5 * g->nx * g->ny * 16;
}
size_t fdfd_factor_bytes(const fdfd_grid *g, int precision) {
  if (check_grid(g) || check_prec(precision)) return 0;
  return factor_total_bytes(g, precision);
}
size_t fdfd_factor_workspace_bytes(const fdfd_grid *g, int precision) {
  if (check_grid(g) || check_prec(precision)) return 0;
  return factor_work_bytes(g, precision);
}
size_t fdfd_solve_workspace_bytes(const fdfd_grid *g, int precision, int nrhs) {
  if (check_grid(g) || check_prec(precision) || nrhs < 1) return 0;
  return solve_work_bytes(g, precision, nrhs);
}

int fdfd_sfactors(const fdfd_grid *g, double *isxf, double *isxb, double *isyf, double *isyb) {
  if (!g || !isxf || !isxb || !isyf || !isyb) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_sfactors");
  if (g->nx < 1 || g->ny < 1 || g->npml_x < 0 || g->npml_y < 0 || !(g->omega > 0) || !(g->dl > 0) || !(g->L0 > 0))
    return fail(FDFD_ERR_ARG, "fdfd_sfactors: bad grid");
  // the same cell sizes and s-values the assembly kernel uses
  const double hx = ((double)g->nx * g->dl) / g->nx, hy = ((double)g->ny * g->dl) / g->ny;
  for (int i = 0; i < g->nx; ++i) {
    const cd f = cinv(sfactor(0, i, g->nx, g->npml_x, hx, g->omega, g->L0));
    const cd b = cinv(sfactor(1, i, g->nx, g->npml_x, hx, g->omega, g->L0));
    isxf[2 * i] = f.r; isxf[2 * i + 1] = f.i;
    isxb[2 * i] = b.r; isxb[2 * i + 1] = b.i;
  }
  for (int j = 0; j < g->ny; ++j) {
    const cd f = cinv(sfactor(0, j, g->ny, g->npml_y, hy, g->omega, g->L0));
    const cd b = cinv(sfactor(1, j, g->ny, g->npml_y, hy, g->omega, g->L0));
    isyf[2 * j] = f.r; isyf[2 * j + 1] = f.i;
    isyb[2 * j] = b.r; isyb[2 * j + 1] = b.i;
  }
  return FDFD_OK;
}

int fdfd_assemble(const fdfd_grid *g, const double *eps, double *op, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (!eps || !op) return fail(FDFD_ERR_ARG, "eps and op must be device pointers");
  hipLaunchKernelGGL(k_assemble, dim3((g->ny + NT - 1) / NT, g->nx), dim3(NT), 0, stream, *g, eps, op);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

int fdfd_flux(int nx, int ny, double dl, int kind, const double *e, const double *h, int i0, int j0, int len,
              double *out, hipStream_t stream) {
  if (!e || !h || !out || kind < 0 || kind > 3 || nx < 1 || ny < 1 || len < 0)
    return fail(FDFD_ERR_ARG, "bad arguments to fdfd_flux");
  const bool xn = kind == 0 || kind == 2;
  // every cell read lies inside the grid: x-normal lines read rows i0, i0+1
  // at columns [j0, j0+len); y-normal lines rows [i0, i0+len) at columns j0, j0+1
  const bool inside = xn ? (i0 >= 0 && i0 + 1 < nx && j0 >= 0 && j0 + len <= ny)
                         : (j0 >= 0 && j0 + 1 < ny && i0 >= 0 && i0 + len <= nx);
  if (!inside)
    return fail(FDFD_ERR_ARG, "fdfd_flux: the probe line and its averaging neighbours must lie inside the grid");
  hipLaunchKernelGGL(k_flux, dim3(1), dim3(NT), 0, stream, ny, dl, kind, e, h, i0, j0, len, out);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

int fdfd_apply(const fdfd_grid *g, const double *op, const double *x, double *y, int nrhs,
               hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (!op || !x || !y || nrhs < 1) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_apply");
  hipLaunchKernelGGL(k_stencil, dim3((g->ny + NT - 1) / NT, g->nx), dim3(NT), 0, stream,
                     g->nx, g->ny, nrhs, op, x, (const double *)nullptr, y);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

int fdfd_residual(const fdfd_grid *g, const double *op, const double *b, const double *x,
                  double *r, int nrhs, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (!op || !x || !b || !r || nrhs < 1) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_residual");
  hipLaunchKernelGGL(k_stencil, dim3((g->ny + NT - 1) / NT, g->nx), dim3(NT), 0, stream,
                     g->nx, g->ny, nrhs, op, x, b, r);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

int fdfd_factor(const fdfd_grid *g, const double *op, void *factor, int precision, void *work,
                hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  if (!op || !factor || !work) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_factor");
  const bool piv = precision & FDFD_FACTOR_PIVOT;
  const bool fp64 = base_prec(precision) == FDFD_PREC_FP64;
  bool ran = false;
  // the persistent kernel factors one chain; a twisted factorization is stepwise
  if (!(precision & FDFD_FACTOR_STEPWISE) && twist_block(g) == g->nx - 1) {
    int rc = fp64 ? (piv ? factor_persistent<double, true>(g, op, factor, work, stream, &ran)
                         : factor_persistent<double, false>(g, op, factor, work, stream, &ran))
                  : (piv ? factor_persistent<float, true>(g, op, factor, work, stream, &ran)
                         : factor_persistent<float, false>(g, op, factor, work, stream, &ran));
    if (rc) return rc;
  }
  if (!ran) {
    int rc = fp64 ? (piv ? factor_impl<double, true>(g, op, factor, work, stream)
                         : factor_impl<double, false>(g, op, factor, work, stream))
                  : (piv ? factor_impl<float, true>(g, op, factor, work, stream)
                         : factor_impl<float, false>(g, op, factor, work, stream));
    if (rc) return rc;
  }
  return fp64 ? wrap_setup<double>(g, op, factor, work, piv, stream)
              : wrap_setup<float>(g, op, factor, work, piv, stream);
}

int fdfd_factor_batch(const fdfd_grid *g, const double *const *ops, void *const *factors, int njobs,
                      int precision, void *const *works, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  if (!ops || !factors || !works || njobs < 1 || njobs > MAX_BATCH)
    return fail(FDFD_ERR_ARG, "bad arguments to fdfd_factor_batch (1 <= njobs <= %d)", MAX_BATCH);
  for (int j = 0; j < njobs; ++j)
    if (!ops[j] || !factors[j] || !works[j]) return fail(FDFD_ERR_ARG, "fdfd_factor_batch: job %d has a NULL pointer", j);
  const bool piv = precision & FDFD_FACTOR_PIVOT;
  const bool fp64 = base_prec(precision) == FDFD_PREC_FP64;
  int rc = fp64 ? (piv ? factor_batch_impl<double, true>(g, ops, factors, works, njobs, stream)
                       : factor_batch_impl<double, false>(g, ops, factors, works, njobs, stream))
                : (piv ? factor_batch_impl<float, true>(g, ops, factors, works, njobs, stream)
                       : factor_batch_impl<float, false>(g, ops, factors, works, njobs, stream));
  if (rc) return rc;
  for (int j = 0; j < njobs; ++j) {
    rc = fp64 ? wrap_setup<double>(g, ops[j], factors[j], works[j], piv, stream)
              : wrap_setup<float>(g, ops[j], factors[j], works[j], piv, stream);
    if (rc) return rc;
  }
  return FDFD_OK;
}
int fdfd_factor_wrap(const fdfd_grid *g, const double *op, void *factor, int precision, void *work,
                     hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  if (!op || !factor || !work) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_factor_wrap");
  const bool piv = precision & FDFD_FACTOR_PIVOT;
  return base_prec(precision) == FDFD_PREC_FP64 ? wrap_setup<double>(g, op, factor, work, piv, stream)
                                                : wrap_setup<float>(g, op, factor, work, piv, stream);
}

int fdfd_factor_status(const fdfd_grid *g, int precision, void *work, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  FactorWork fw = factor_work_layout(g, precision, work);
  int st = 0;
  HIP_TRY(hipMemcpyAsync(&st, fw.status, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (st == 2) return fail(FDFD_ERR_NUMERIC, "a wait inside a block factorization step timed out");
  if (st) return fail(FDFD_ERR_NUMERIC, "zero or non-finite pivot in the block factorization");
  return FDFD_OK;
}

int fdfd_sweep(const fdfd_grid *g, const double *op, const void *factor, int precision,
               const double *r, double *x, int nrhs, void *work, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  if (!op || !factor || !r || !x || !work || nrhs < 1) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_sweep");
  SolveWork sw = solve_work_layout(g, precision, nrhs, work);
  HIP_TRY(hipMemsetAsync(sw.sync, 0, sizeof(SweepSync), stream));
  int rc = base_prec(precision) == FDFD_PREC_FP64
               ? sweep_impl<double>(g, op, factor, r, x, nrhs, sw, stream, false)
               : sweep_impl<float>(g, op, factor, r, x, nrhs, sw, stream, false);
  if (rc) return rc;
  unsigned ab = 0;
  HIP_TRY(hipMemcpyAsync(&ab, &sw.sync->abort, sizeof ab, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (ab) return fail(FDFD_ERR_HIP, "sweep workgroups were not co-resident; x is invalid");
  return FDFD_OK;
}

int fdfd_norms(const fdfd_grid *g, const double *v, int nrhs, void *work, double *out,
               hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (!v || !work || !out || nrhs < 1) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_norms");
  SolveWork sw = solve_work_layout(g, FDFD_PREC_FP64, nrhs, work);
  if (int rc = norms_async(g, v, nrhs, 0, sw.part, sw.norms, stream)) return rc;
  HIP_TRY(hipMemcpyAsync(out, sw.norms, sizeof(double) * nrhs, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  return FDFD_OK;
}

int fdfd_stretch_scale(const fdfd_grid *g, double *v, int nrhs, int inverse, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (!v || nrhs < 1) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_stretch_scale");
  hipLaunchKernelGGL(k_stretch_scale, dim3((g->ny + NT - 1) / NT, g->nx, nrhs), dim3(NT), 0, stream,
                     *g, v, inverse);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

}  // extern "C"

namespace {
// The refinement of fdfd_solve. work: the solve workspace (device); host:
// 4 * nrhs doubles of host scratch (norms of b, of the residual, the relative
// residuals and the previous step's).
int solve_impl(const fdfd_grid *g, const double *op, const void *factor, int precision, const double *b,
               double *x, int nrhs, int transpose, double tol, int max_iter, void *work,
               fdfd_solve_stats *stats, double *rel_res, hipStream_t stream, double *host) {
  SolveWork sw = solve_work_layout(g, precision, nrhs, work);
  const size_t M = (size_t)g->nx * g->ny;
  const dim3 sgrid((g->ny + NT - 1) / NT, g->nx);
  fdfd_solve_stats st;
  memset(&st, 0, sizeof st);
  double *nb = host, *nr = host + nrhs, *rel = host + 2 * nrhs, *prev = host + 3 * nrhs;
  for (int q = 0; q < nrhs; ++q) rel[q] = 0.0;
  // ||b|| in the caller's system
  if (int rc = norms_async(g, b, nrhs, 0, sw.part, sw.norms, stream)) return rc;
  HIP_TRY(hipMemcpyAsync(nb, sw.norms, sizeof(double) * nrhs, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  bool all_zero = true;
  for (int q = 0; q < nrhs; ++q) all_zero = all_zero && nb[q] == 0.0;
  if (all_zero) {  // linalg.py:128-129
    HIP_TRY(hipMemsetAsync(x, 0, 16 * M * nrhs, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    st.zero_rhs = 1;
    st.converged = 1;
    if (stats) *stats = st;
    if (rel_res) for (int q = 0; q < nrhs; ++q) rel_res[q] = 0;
    return FDFD_OK;
  }
  // transpose: A^T = D A D^-1, so A^T x = b is A x' = D^-1 b with x = D x'
  double *rr = sw.r;  // residual of the iterate, in the solved system
  int rc = FDFD_OK;
  bool blockwise = false;
  for (int attempt = 0; attempt < 2; ++attempt) {
    memset(&st, 0, sizeof st);
    HIP_TRY(hipMemsetAsync(sw.sync, 0, sizeof(SweepSync), stream));
    HIP_TRY(hipMemsetAsync(x, 0, 16 * M * nrhs, stream));
    HIP_TRY(hipMemcpyAsync(rr, b, 16 * M * nrhs, hipMemcpyDeviceToDevice, stream));
    if (transpose && (rc = fdfd_stretch_scale(g, rr, nrhs, 1, stream))) return rc;
    for (int q = 0; q < nrhs; ++q) prev[q] = INFINITY;
    int stall_count = 0;
    bool aborted = false;
    const bool fp64 = base_prec(precision) == FDFD_PREC_FP64;
    const double *kinv = wrap_kinv(g, precision, factor);
    const bool exact = wrap_exact(g);
    for (int it = 1; it <= max_iter; ++it) {
      HIP_TRY(hipMemsetAsync(sw.t, 0, 16 * M * nrhs, stream));
      if (exact) {
        // t = T^-1 r; f = -U C K^-1 V^T t; t += T^-1 f; x += t  (x += A^-1 r)
        rc = fp64 ? sweep_impl<double>(g, op, factor, rr, sw.t, nrhs, sw, stream, blockwise)
                  : sweep_impl<float>(g, op, factor, rr, sw.t, nrhs, sw, stream, blockwise);
        if (rc) return rc;
        HIP_TRY(hipMemsetAsync(sw.f, 0, 16 * M * nrhs, stream));
        hipLaunchKernelGGL(k_wrap_rhs, dim3((2 * g->ny + NT / 64 - 1) / (NT / 64), nrhs), dim3(NT), 0, stream, *g,
                           op, kinv, (const double *)sw.t, sw.f);
        HIP_TRY(hipGetLastError());
        rc = fp64 ? sweep_impl<double>(g, op, factor, sw.f, sw.t, nrhs, sw, stream, blockwise)
                  : sweep_impl<float>(g, op, factor, sw.f, sw.t, nrhs, sw, stream, blockwise);
        if (rc) return rc;
        st.sweeps += 2;
      } else {
        // One sweep per step. The previous step left r = g + d, g on the
        // boundary blocks (its uncorrected wrap coupling) and d small. The
        // forcing f corrects g exactly (truncation aside): x += T^-1 (r + f),
        // and the new boundary part g' follows from f and the sweep's
        // boundary blocks. The first step has no g.
        if (it > 1 && (rc = fp64 ? wrap_lag_forcing<double>(g, op, factor, sw, nrhs, rr, stream)
                                 : wrap_lag_forcing<float>(g, op, factor, sw, nrhs, rr, stream)))
          return rc;
        rc = fp64 ? sweep_impl<double>(g, op, factor, rr, sw.t, nrhs, sw, stream, blockwise)
                  : sweep_impl<float>(g, op, factor, rr, sw.t, nrhs, sw, stream, blockwise);
        if (rc) return rc;
        hipLaunchKernelGGL(k_wrap_lag_next, dim3((g->ny + NT - 1) / NT, nrhs), dim3(NT), 0, stream, *g, op,
                           (const double *)sw.t, it > 1 ? 1 : 0, sw.f);
        HIP_TRY(hipGetLastError());
        st.sweeps += 1;
      }
      hipLaunchKernelGGL(k_axpy, dim3(1024), dim3(NT), 0, stream, 2 * M * nrhs, (const double *)sw.t, x);
      HIP_TRY(hipGetLastError());
      st.iterations = it;
      // r = b' - A x, with b' = b or D^-1 b
      HIP_TRY(hipMemcpyAsync(rr, b, 16 * M * nrhs, hipMemcpyDeviceToDevice, stream));
      if (transpose && (rc = fdfd_stretch_scale(g, rr, nrhs, 1, stream))) return rc;
      hipLaunchKernelGGL(k_stencil, sgrid, dim3(NT), 0, stream, g->nx, g->ny, nrhs, op, (const double *)x,
                         (const double *)rr, rr);
      HIP_TRY(hipGetLastError());
      // transposed: ||b - A^T x|| = ||D (b' - A x')||
      if ((rc = norms_async(g, rr, nrhs, transpose, sw.part, sw.norms, stream))) return rc;
      unsigned ab = 0;
      HIP_TRY(hipMemcpyAsync(nr, sw.norms, sizeof(double) * nrhs, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipMemcpyAsync(&ab, &sw.sync->abort, sizeof ab, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      if (ab && !blockwise) { aborted = true; break; }
      double worst = 0;
      bool contracting = false;
      for (int q = 0; q < nrhs; ++q) {
        rel[q] = nb[q] > 0 ? sqrt(nr[q] / nb[q]) : 0.0;
        if (!std::isfinite(rel[q])) rel[q] = INFINITY;
        worst = std::max(worst, rel[q]);
        if (rel[q] > tol && rel[q] < 0.7 * prev[q]) contracting = true;
      }
      st.max_rel_residual = worst;
      if (worst <= tol) { st.converged = 1; break; }
      if (!std::isfinite(worst)) { st.stalled = 1; break; }
      stall_count = contracting ? 0 : stall_count + 1;
      if (stall_count >= 2) { st.stalled = 1; break; }
      for (int q = 0; q < nrhs; ++q) prev[q] = rel[q];
    }
    if (!aborted) break;
    blockwise = true;  // persistent sweep could not run co-resident: redo with per-block launches
  }
  st.restarts = blockwise ? 1 : 0;
  if (transpose) {
    if ((rc = fdfd_stretch_scale(g, x, nrhs, 0, stream))) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
  }
  if (stats) *stats = st;
  if (rel_res) for (int q = 0; q < nrhs; ++q) rel_res[q] = rel[q];
  if (!st.converged)
    return fail(FDFD_ERR_NOT_CONVERGED, "refinement reached %.3e after %d sweeps (tol %.1e)%s",
                st.max_rel_residual, st.iterations, tol, st.stalled ? ", stalled" : "");
  return FDFD_OK;
}

// ------------------------------------------- solve without allocations
// The workspace of fdfd_solve_static: one array per device, at a fixed
// address for the life of the process. Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device (the array's
// name is a host-side stand-in, not a device address).
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int STATIC_MAX_RHS = 16;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// this device's copy of scratch (looked up once per device)
int static_scratch(void **ptr) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(FDFD_ERR_ARG, "device %d beyond the supported %d", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(FDFD_ERR_HIP, "hipGetSymbolAddress returned NULL for the solve scratch");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return FDFD_OK;
}
}  // namespace

extern "C" {

int fdfd_solve(const fdfd_grid *g, const double *op, const void *factor, int precision,
               const double *b, double *x, int nrhs, int transpose, double tol, int max_iter,
               void *work, fdfd_solve_stats *stats, double *rel_res, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  if (!op || !factor || !b || !x || !work || nrhs < 1 || max_iter < 1 || !(tol > 0))
    return fail(FDFD_ERR_ARG, "bad arguments to fdfd_solve");
  std::vector<double> host(4 * (size_t)nrhs);
  return solve_impl(g, op, factor, precision, b, x, nrhs, transpose, tol, max_iter, work, stats, rel_res,
                    stream, host.data());
}

int fdfd_solve_static(const fdfd_grid *g, const double *op, const void *factor, int precision,
                      const double *b, double *x, int nrhs, int transpose, double tol, int max_iter,
                      fdfd_solve_stats *stats, double *rel_res, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  if (!op || !factor || !b || !x || nrhs < 1 || max_iter < 1 || !(tol > 0))
    return fail(FDFD_ERR_ARG, "bad arguments to fdfd_solve_static");
  if (nrhs > STATIC_MAX_RHS)
    return fail(FDFD_ERR_ARG, "fdfd_solve_static: %d right-hand sides exceed the supported %d", nrhs,
                STATIC_MAX_RHS);
  const size_t need = solve_work_bytes(g, precision, nrhs);
  if (need > STATIC_SCRATCH_BYTES)
    return fail(FDFD_ERR_ARG, "fdfd_solve_static: the solve needs %zu bytes of workspace, beyond the static "
                "%zu (use fdfd_solve with a caller workspace)", need, STATIC_SCRATCH_BYTES);
  void *work = nullptr;
  if (int rc = static_scratch(&work)) return rc;
  double host[4 * STATIC_MAX_RHS];
  return solve_impl(g, op, factor, precision, b, x, nrhs, transpose, tol, max_iter, work, stats, rel_res,
                    stream, host);
}

int fdfd_grad_eps(const fdfd_grid *g, const double *lam, const double *e, double *grad,
                  hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (g->pol != FDFD_POL_EZ) return fail(FDFD_ERR_ARG, "fdfd_grad_eps supports pol='Ez' only");
  if (!lam || !e || !grad) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_grad_eps");
  const size_t M = (size_t)g->nx * g->ny;
  hipLaunchKernelGGL(k_grad_eps, dim3((M + NT - 1) / NT), dim3(NT), 0, stream, *g, lam, e, grad);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

int fdfd_kerr_update(const fdfd_grid *g, double *op, const double *base_c, const double *e,
                     const double *chi_region, double *eps_nl_out, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (g->pol != FDFD_POL_EZ) return fail(FDFD_ERR_ARG, "Kerr update supports pol='Ez' only");
  if (!op || !base_c || !e || !chi_region) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_kerr_update");
  const size_t M = (size_t)g->nx * g->ny;
  hipLaunchKernelGGL(k_kerr, dim3((M + NT - 1) / NT), dim3(NT), 0, stream, *g, op, base_c, e,
                     chi_region, eps_nl_out);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

int fdfd_fields(const fdfd_grid *g, const double *eps, const double *x, double *f1, double *f2,
                int nrhs, int averaging, hipStream_t stream) {
  if (int rc = check_grid(g)) return rc;
  if (!x || !f1 || !f2 || nrhs < 1 || (g->pol == FDFD_POL_HZ && !eps))
    return fail(FDFD_ERR_ARG, "bad arguments to fdfd_fields");
  hipLaunchKernelGGL(k_fields, dim3((g->ny + NT - 1) / NT, g->nx, nrhs), dim3(NT), 0, stream, *g,
                     eps, x, f1, f2, averaging);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

fdfd_engine *fdfd_engine_create(int device) {
  auto *E = new (std::nothrow) fdfd_engine();
  if (!E) { fail(FDFD_ERR_HIP, "out of host memory"); return nullptr; }
  E->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
    fail(FDFD_ERR_HIP, "cannot create the engine stream on device %d", device);
    delete E;
    return nullptr;
  }
  E->worker = std::thread(engine_loop, E);
  return E;
}

int fdfd_engine_destroy(fdfd_engine *E) {
  if (!E) return FDFD_OK;
  {
    std::lock_guard<std::mutex> lk(E->mu);
    E->stop = true;
  }
  E->cv_work.notify_all();
  E->worker.join();
  (void)hipStreamSynchronize(E->stream);
  for (auto &kv : E->jobs)
    if (kv.second->done) (void)hipEventDestroy(kv.second->done);
  (void)hipStreamDestroy(E->stream);
  delete E;
  return FDFD_OK;
}

int fdfd_engine_submit(fdfd_engine *E, const fdfd_grid *g, const double *op, void *factor, int precision,
                       void *work, uint64_t *ticket) {
  if (!E || !ticket) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_engine_submit");
  if (int rc = check_grid(g)) return rc;
  if (int rc = check_prec(precision)) return rc;
  if (!op || !factor || !work) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_engine_submit");
  auto e = std::make_shared<EngineJob>();
  e->g = *g; e->op = op; e->factor = factor; e->work = work;
  e->precision = base_prec(precision); e->pivot = precision & FDFD_FACTOR_PIVOT;
  e->m = g->nx; e->n = g->ny; e->np = pad_n(g->ny); e->nt = e->np / TB;
  e->sym = g_step_sym.load() != 0;
  {
    const int c = twist_block(g);
    e->nch = c < e->m - 1 ? 2 : 1;
    e->steps = (c + 1) * e->nt;
  }
  HIP_TRY(hipEventCreateWithFlags(&e->done, hipEventDisableTiming));
  {
    std::lock_guard<std::mutex> lk(E->mu);
    *ticket = E->next_ticket++;
    E->jobs[*ticket] = e;
    E->pending.push_back(e);
  }
  E->cv_work.notify_all();
  return FDFD_OK;
}

int fdfd_engine_wait(fdfd_engine *E, uint64_t ticket) {
  if (!E) return fail(FDFD_ERR_ARG, "engine is NULL");
  std::shared_ptr<EngineJob> e;
  int rc = FDFD_OK;
  {
    std::unique_lock<std::mutex> lk(E->mu);
    auto it = E->jobs.find(ticket);
    if (it == E->jobs.end()) return fail(FDFD_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    e = it->second;
    E->cv_done.wait(lk, [&] { return e->launched_all; });
    rc = e->rc;
    E->jobs.erase(it);
  }
  // rc == FDFD_OK: the event was recorded before launched_all was set
  hipError_t err = rc == FDFD_OK ? hipEventSynchronize(e->done) : hipSuccess;
  if (rc != FDFD_OK) (void)hipStreamSynchronize(E->stream);  // no step of the job still queued
  (void)hipEventDestroy(e->done);
  e->done = nullptr;
  if (err != hipSuccess) return fail(FDFD_ERR_HIP, "engine factor: %s", hipGetErrorString(err));
  if (rc) return fail(rc, "engine factor launch failed");
  return FDFD_OK;
}

uint64_t fdfd_engine_launches(fdfd_engine *E) {
  if (!E) return 0;
  std::lock_guard<std::mutex> lk(E->mu);
  return E->launches;
}

void fdfd_set_spin_cap(unsigned polls) { g_spin_cap.store(polls ? polls : SPIN_CAP); }

int fdfd_set_step_symmetry(int on) {
  const int prev = g_step_sym.exchange(on ? 1 : 0);
  return prev;
}

int fdfd_set_twist(int on) { return g_twist.exchange(on ? 1 : 0); }

int fdfd_twist_block(const fdfd_grid *g) {
  if (check_grid(g)) return -1;
  return twist_block(g);
}

int fdfd_selftest_mfma(int precision, const double *a, const double *b, double *d, hipStream_t stream) {
  if (precision < 0 || precision > 2) return fail(FDFD_ERR_ARG, "selftest mode must be 0, 1 or 2");
  if (!a || !b || !d) return fail(FDFD_ERR_ARG, "bad arguments to fdfd_selftest_mfma");
  if (precision == FDFD_PREC_FP64)
    hipLaunchKernelGGL(k_selftest_mfma_f64, dim3(1), dim3(64), 0, stream, a, b, d, 0);
  else if (precision == 2)
    hipLaunchKernelGGL(k_selftest_mfma_f64, dim3(1), dim3(64), 0, stream, a, b, d, 1);
  else
    hipLaunchKernelGGL(k_selftest_mfma_f32, dim3(1), dim3(64), 0, stream, a, b, d);
  HIP_TRY(hipGetLastError());
  return FDFD_OK;
}

}  // extern "C"
