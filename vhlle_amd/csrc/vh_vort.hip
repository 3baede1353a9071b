// Resolution of the vorticity matrix entries that the reference inherits from
// earlier cells (see vh_vort.h): four launches over the (ix, iy) columns of a
// single grid, M = nx * ny columns, loop-order column index K = iy + ny * ix.
//
//  k_col    one thread per column walks iz upwards (adjacent threads read
//           adjacent cells) carrying the matrix from the column's own set
//           entries; records per group whether the column sets it, whether a
//           cell needs the value that enters the column, and the value that
//           leaves it.
//  k_scan   one block: for every column and group, the last earlier column
//           that changes the group (prefix max) and, for [3][0]/[3][3], the
//           next one (suffix min).
//  k_chain  a column that only ADDS to [3][0]/[3][3] passes its input on plus
//           its adds: each chain of such columns is folded by one thread, in
//           order, from the column that set the entries before it.
//  k_fill   columns with cells that need the entering value walk iz again up
//           to each group's first set.
#include "vh_vort.h"

namespace vh {
namespace vort {
namespace {

constexpr int INFO_NEED = 5;      // bits 0-4: the column sets group g; 5-9: a cell needs its input
constexpr int INFO_ADDONLY = 10;  // [3][0], [3][3] added to but never set in the column
constexpr int SCAN_THREADS = 1024;
constexpr int PREF = 8;           // op codes fetched ahead per column walk

struct Scratch {
  int *info, *prev, *next4;
  double *out;  // [K][16]: the matrix leaving column K (groups it changes)
};

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

Scratch carve(void *p, long M) {
  char *b = static_cast<char *>(p);
  Scratch s;
  s.info = reinterpret_cast<int *>(b);
  b += al256(M * sizeof(int));
  s.prev = reinterpret_cast<int *>(b);
  b += al256(NGROUP * M * sizeof(int));
  s.next4 = reinterpret_cast<int *>(b);
  b += al256(M * sizeof(int));
  s.out = reinterpret_cast<double *>(b);
  return s;
}

__device__ inline bool changes(int info, int gi) {
  return ((info >> gi) & 1) || (gi == 4 && ((info >> INFO_ADDONLY) & 1));
}

__global__ void __launch_bounds__(256) k_col(Geom g, double *__restrict__ db,
                                             const unsigned char *__restrict__ op, Scratch s) {
  const long M = (long)g.nx * g.ny, N = g.N;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M) return;
  const int ix = (int)(t % g.nx), iy = (int)(t / g.nx);
  const long K = iy + (long)g.ny * ix;
  long ls[NGROUP];   // last cell of the column that set the group (-1: none yet)
  bool ld[NGROUP];   // cv holds that cell's (possibly since added-to) values
  double cv[16];
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) {
    ls[gi] = -1;
    ld[gi] = false;
  }
#pragma unroll
  for (int k = 0; k < 16; k++) cv[k] = 0.;
  int need = 0, addunk = 0;
  for (int z0 = 0; z0 < g.nz; z0 += PREF) {
    unsigned long long ob = 0;  // PREF op codes, one byte each
#pragma unroll
    for (int j = 0; j < PREF; j++)
      if (z0 + j < g.nz) ob |= (unsigned long long)op[t + M * (z0 + j)] << (8 * j);
#pragma unroll 1
    for (int j = 0; j < PREF; j++) {
      const unsigned char o = (unsigned char)(ob >> (8 * j));
      if ((o & K_MASK) == K_EMPTY) continue;
      const long c = t + M * (z0 + j);
      if (self_contained(o)) {
#pragma unroll
        for (int gi = 0; gi < NGROUP; gi++) {
          ls[gi] = c;
          ld[gi] = false;
        }
        continue;
      }
#pragma unroll
      for (int gi = 0; gi < NGROUP; gi++) {
        const int gop = group_op(o, gi);
        if (gop == G_SET) {
          ls[gi] = c;
          ld[gi] = false;
          continue;
        }
        if (ls[gi] < 0) {  // depends on the matrix entering the column
          need |= 1 << gi;
          if (gop == G_ADD) addunk = 1;
          continue;
        }
        if (!ld[gi]) {
#pragma unroll
          for (int k = 0; k < group_size(gi); k++) {
            const int e = group_entry(gi, k);
            cv[e] = db[e * N + ls[gi]];
          }
          ld[gi] = true;
        }
#pragma unroll
        for (int k = 0; k < group_size(gi); k++) {
          const int e = group_entry(gi, k);
          if (gop == G_ADD) cv[e] += db[e * N + c];
          db[e * N + c] = cv[e];
        }
      }
    }
  }
  int info = need << INFO_NEED;
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) {
    if (ls[gi] < 0) continue;
    info |= 1 << gi;
#pragma unroll
    for (int k = 0; k < group_size(gi); k++) {
      const int e = group_entry(gi, k);
      s.out[K * 16 + e] = ld[gi] ? cv[e] : db[e * N + ls[gi]];
    }
  }
  if (addunk && ls[4] < 0) info |= 1 << INFO_ADDONLY;
  s.info[K] = info;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan(int M, Scratch s) {
  __shared__ int lastv[NGROUP][SCAN_THREADS];
  __shared__ int first4[SCAN_THREADS];
  const int i = threadIdx.x;
  const int ch = (M + SCAN_THREADS - 1) / SCAN_THREADS;
  const int k0 = i * ch < M ? i * ch : M;
  const int k1 = k0 + ch < M ? k0 + ch : M;
  int last[NGROUP], f4 = M;
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) last[gi] = -1;
  for (int K = k0; K < k1; K++) {
    const int inf = s.info[K];
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++)
      if (changes(inf, gi)) last[gi] = K;
    if (f4 == M && changes(inf, 4)) f4 = K;
  }
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) lastv[gi][i] = last[gi];
  first4[i] = f4;
  __syncthreads();
  // inclusive prefix max (lastv) and suffix min (first4) over the threads
  for (int d = 1; d < SCAN_THREADS; d <<= 1) {
    int a[NGROUP];
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) a[gi] = i >= d ? lastv[gi][i - d] : -1;
    const int b = i + d < SCAN_THREADS ? first4[i + d] : M;
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) lastv[gi][i] = lastv[gi][i] > a[gi] ? lastv[gi][i] : a[gi];
    first4[i] = first4[i] < b ? first4[i] : b;
    __syncthreads();
  }
  int run[NGROUP];
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) run[gi] = i > 0 ? lastv[gi][i - 1] : -1;
  for (int K = k0; K < k1; K++) {
    const int inf = s.info[K];
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) {
      s.prev[(long)gi * M + K] = run[gi];
      if (changes(inf, gi)) run[gi] = K;
    }
  }
  int nx4 = i + 1 < SCAN_THREADS ? first4[i + 1] : M;
  for (int K = k1 - 1; K >= k0; K--) {
    s.next4[K] = nx4;
    if (changes(s.info[K], 4)) nx4 = K;
  }
}

__global__ void __launch_bounds__(256) k_chain(Geom g, const double *__restrict__ db,
                                               const unsigned char *__restrict__ op, Scratch s) {
  const int M = g.nx * g.ny;
  const long N = g.N;
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= M) return;
  if (!((s.info[K] >> INFO_ADDONLY) & 1)) return;
  const int p = s.prev[4L * M + K];
  if (p >= 0 && !((s.info[p] >> 4) & 1)) return;  // K continues a chain, not its head
  double v0 = 0., v3 = 0.;
  if (p >= 0) {
    v0 = s.out[(long)p * 16 + 12];
    v3 = s.out[(long)p * 16 + 15];
  }
  int P = K;
  for (;;) {
    const long t = (long)(P / g.ny) + (long)g.nx * (P % g.ny);
    for (int iz = 0; iz < g.nz; iz++) {
      const long c = t + (long)M * iz;
      if (group_op(op[c], 4) == G_ADD) {
        v0 += db[12 * N + c];
        v3 += db[15 * N + c];
      }
    }
    s.out[(long)P * 16 + 12] = v0;
    s.out[(long)P * 16 + 15] = v3;
    const int q = s.next4[P];  // > P, or M
    if (q >= M || !((s.info[q] >> INFO_ADDONLY) & 1)) break;
    P = q;
  }
}

__global__ void __launch_bounds__(256) k_fill(Geom g, double *__restrict__ db,
                                              const unsigned char *__restrict__ op, Scratch s) {
  const long M = (long)g.nx * g.ny, N = g.N;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M) return;
  const int ix = (int)(t % g.nx), iy = (int)(t / g.nx);
  const long K = iy + (long)g.ny * ix;
  int pend = (s.info[K] >> INFO_NEED) & 31;
  if (!pend) return;
  double cv[16];
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) {
    const int p = (pend >> gi) & 1 ? s.prev[gi * M + K] : -1;
#pragma unroll
    for (int k = 0; k < group_size(gi); k++) {
      const int e = group_entry(gi, k);
      cv[e] = p >= 0 ? s.out[(long)p * 16 + e] : 0.;
    }
  }
  for (int iz = 0; iz < g.nz && pend; iz++) {
    const long c = t + M * iz;
    const unsigned char o = op[c];
    if ((o & K_MASK) == K_EMPTY) continue;
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) {
      if (!((pend >> gi) & 1)) continue;
      const int gop = group_op(o, gi);
      if (gop == G_SET) {  // from here on the column's own walk wrote the group
        pend &= ~(1 << gi);
        continue;
      }
#pragma unroll
      for (int k = 0; k < group_size(gi); k++) {
        const int e = group_entry(gi, k);
        if (gop == G_ADD) cv[e] += db[e * N + c];
        db[e * N + c] = cv[e];
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_clamp(double *__restrict__ a, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = clamp100(a[i]);
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace vort

size_t vort_scratch_bytes(const Geom &g) {
  const long M = (long)g.nx * g.ny;
  return vort::al256(M * sizeof(int)) + vort::al256(vort::NGROUP * M * sizeof(int)) +
         vort::al256(M * sizeof(int)) + vort::al256(M * 16 * sizeof(double));
}

void vort_resolve(const Geom &g, double *db, const unsigned char *op, void *scratch,
                  hipStream_t st) {
  using namespace vort;
  const long M = (long)g.nx * g.ny;
  const Scratch s = carve(scratch, M);
  hipLaunchKernelGGL(k_col, dim3(nblk(M)), dim3(256), 0, st, g, db, op, s);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_THREADS), 0, st, (int)M, s);
  hipLaunchKernelGGL(k_chain, dim3(nblk(M)), dim3(256), 0, st, g, db, op, s);
  hipLaunchKernelGGL(k_fill, dim3(nblk(M)), dim3(256), 0, st, g, db, op, s);
}

void vort_clamp(double *db, long n, hipStream_t st) {
  hipLaunchKernelGGL(vort::k_clamp, dim3(vort::nblk(n)), dim3(256), 0, st, db, n);
}

}  // namespace vh
