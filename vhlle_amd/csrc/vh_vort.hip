// Resolution of the vorticity matrix entries that the reference inherits from
// earlier cells (see vh_vort.h): four launches over the (ix, iy) columns of a
// single grid, M = nx * ny columns, loop-order column index K = iy + ny * ix.
//
//  k_col    one thread per column walks iz upwards (adjacent threads read
//           adjacent cells) carrying the matrix from the column's own set
//           entries; records per group whether the column sets it, whether a
//           cell needs the value that enters the column, and the value that
//           leaves it.
//  k_tile, k_carry, k_tile
//           for every column and group, the last earlier column that changes
//           the group (prefix max) and, for [3][0]/[3][3], the next one
//           (suffix min): scans within tiles of 1024 columns, a scan of the
//           tiles' aggregates, then the tiles again with their carries.
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
constexpr int TILE = 1024;         // columns per scan tile (one thread each)
constexpr int PREF = 8;           // op codes fetched ahead per column walk

struct Scratch {
  int *info, *prev, *next4;
  int *agg, *carry;  // [NGROUP + 1][tiles]: per tile last changing columns and first
                     // [3][0]/[3][3] change; then the carries from the other tiles
  double *out;       // [K][16]: the matrix leaving column K (groups it changes)
};

inline long ntiles(long M) { return (M + TILE - 1) / TILE; }

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
  s.agg = reinterpret_cast<int *>(b);
  b += al256((NGROUP + 1) * ntiles(M) * sizeof(int));
  s.carry = reinterpret_cast<int *>(b);
  b += al256((NGROUP + 1) * ntiles(M) * sizeof(int));
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

// inclusive prefix max of pm[g][.] and suffix min of sm[.] over the TILE
// threads of the block (LDS, Hillis-Steele); every thread calls it
__device__ inline void block_scan(int (*pm)[TILE], int *sm, int M) {
  const int i = threadIdx.x;
  __syncthreads();
  for (int d = 1; d < TILE; d <<= 1) {
    int a[NGROUP];
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) a[gi] = i >= d ? pm[gi][i - d] : -1;
    const int b = i + d < TILE ? sm[i + d] : M;
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) pm[gi][i] = pm[gi][i] > a[gi] ? pm[gi][i] : a[gi];
    sm[i] = sm[i] < b ? sm[i] : b;
    __syncthreads();
  }
}

// one thread per column K of tile blockIdx.x.  PASS 0: the tile's aggregates
// (its last column changing each group, its first changing [3][0]/[3][3]).
// PASS 1: prev and next4 of every column from the tile's scans and carries.
template <int PASS>
__global__ void __launch_bounds__(TILE) k_tile(int M, Scratch s) {
  __shared__ int pm[NGROUP][TILE];
  __shared__ int sm[TILE];
  const int i = threadIdx.x, b = blockIdx.x, nb = gridDim.x;
  const int K = b * TILE + i;
  const int inf = K < M ? s.info[K] : 0;
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) pm[gi][i] = K < M && changes(inf, gi) ? K : -1;
  sm[i] = K < M && changes(inf, 4) ? K : M;
  block_scan(pm, sm, M);
  if (PASS == 0) {
    if (i == TILE - 1)
#pragma unroll
      for (int gi = 0; gi < NGROUP; gi++) s.agg[gi * nb + b] = pm[gi][i];
    if (i == 0) s.agg[NGROUP * nb + b] = sm[0];
    return;
  }
  if (K >= M) return;
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) {
    int p = s.carry[gi * nb + b];
    if (i > 0 && pm[gi][i - 1] > p) p = pm[gi][i - 1];
    s.prev[(long)gi * M + K] = p;
  }
  int q = s.carry[NGROUP * nb + b];
  if (i + 1 < TILE && sm[i + 1] < q) q = sm[i + 1];
  s.next4[K] = q;
}

// one block over the nb tile aggregates: carry[g][b] = the last column
// changing group g in the tiles before b (-1: none), carry[NGROUP][b] = the
// first column changing [3][0]/[3][3] in the tiles after b (M: none)
__global__ void __launch_bounds__(TILE) k_carry(int nb, int M, Scratch s) {
  __shared__ int pm[NGROUP][TILE];
  __shared__ int sm[TILE];
  const int i = threadIdx.x;
  const int ch = (nb + TILE - 1) / TILE;
  const int b0 = i * ch < nb ? i * ch : nb;
  const int b1 = b0 + ch < nb ? b0 + ch : nb;
  int last[NGROUP], f = M;
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) last[gi] = -1;
  for (int b = b0; b < b1; b++)
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) {
      const int a = s.agg[gi * nb + b];
      last[gi] = a > last[gi] ? a : last[gi];
    }
  for (int b = b0; b < b1; b++) {
    const int a = s.agg[NGROUP * nb + b];
    f = a < f ? a : f;
  }
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) pm[gi][i] = last[gi];
  sm[i] = f;
  block_scan(pm, sm, M);
  int run[NGROUP];
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) run[gi] = i > 0 ? pm[gi][i - 1] : -1;
  for (int b = b0; b < b1; b++)
#pragma unroll
    for (int gi = 0; gi < NGROUP; gi++) {
      s.carry[gi * nb + b] = run[gi];
      const int a = s.agg[gi * nb + b];
      run[gi] = a > run[gi] ? a : run[gi];
    }
  int nf = i + 1 < TILE ? sm[i + 1] : M;
  for (int b = b1 - 1; b >= b0; b--) {
    s.carry[NGROUP * nb + b] = nf;
    const int a = s.agg[NGROUP * nb + b];
    nf = a < nf ? a : nf;
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
  const size_t aggb = vort::al256((vort::NGROUP + 1) * vort::ntiles(M) * sizeof(int));
  return vort::al256(M * sizeof(int)) + vort::al256(vort::NGROUP * M * sizeof(int)) +
         vort::al256(M * sizeof(int)) + 2 * aggb + vort::al256(M * 16 * sizeof(double));
}

void vort_resolve(const Geom &g, double *db, const unsigned char *op, void *scratch,
                  hipStream_t st) {
  using namespace vort;
  const long M = (long)g.nx * g.ny;
  const Scratch s = carve(scratch, M);
  hipLaunchKernelGGL(k_col, dim3(nblk(M)), dim3(256), 0, st, g, db, op, s);
  const unsigned nb = (unsigned)ntiles(M);
  hipLaunchKernelGGL(k_tile<0>, dim3(nb), dim3(TILE), 0, st, (int)M, s);
  hipLaunchKernelGGL(k_carry, dim3(1), dim3(TILE), 0, st, (int)nb, (int)M, s);
  hipLaunchKernelGGL(k_tile<1>, dim3(nb), dim3(TILE), 0, st, (int)M, s);
  hipLaunchKernelGGL(k_chain, dim3(nblk(M)), dim3(256), 0, st, g, db, op, s);
  hipLaunchKernelGGL(k_fill, dim3(nblk(M)), dim3(256), 0, st, g, db, op, s);
}

void vort_clamp(double *db, long n, hipStream_t st) {
  hipLaunchKernelGGL(vort::k_clamp, dim3(vort::nblk(n)), dim3(256), 0, st, db, n);
}

}  // namespace vh
