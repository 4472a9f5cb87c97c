// Sparse assembly: contact records + static elastic blocks -> the reference's CSR.
//
// Reference: BHessian::toTriplets (CPU IPC/ContactMechanics.cpp:13-134) and
// NewtonLinearSystem::assemble (NewtonLinearSystem.cpp:126-154): lower-triangle entries
// (row >= col) of blocks whose two vertices are free, entries exactly equal to zero
// dropped before duplicates are summed, the mass diagonal on every DOF, Eigen's
// column-major lower storage (outer = column starts, inner = ascending rows), and the
// right-hand side zeroed on non-free vertices.
//
// Pipeline: sort contact block records by (column vertex, row vertex), sum equal keys
// in record order; merge them per block column with the static pattern; expand the
// merged 3x3 blocks to scalar entries keeping exactly those that received a nonzero
// contribution.
#include "context.hpp"

#include <hipcub/hipcub.hpp>

namespace cipc {

constexpr unsigned long long kNoKey = ~0ull;

void sort_u64_pairs(Context& c, const unsigned long long* kin, unsigned long long* kout, const int* vin, int* vout,
                    int n, int bits) {
    if (n <= 0) return;
    size_t bytes = 0;
    CIPC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, n, 0, bits, c.stream));
    c.tmp.reserve(bytes);
    CIPC_HIP(hipcub::DeviceRadixSort::SortPairs(c.tmp.p, bytes, kin, kout, vin, vout, n, 0, bits, c.stream));
}
void sort_u64(Context& c, const unsigned long long* kin, unsigned long long* kout, int n, int bits) {
    if (n <= 0) return;
    size_t bytes = 0;
    CIPC_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, kin, kout, n, 0, bits, c.stream));
    c.tmp.reserve(bytes);
    CIPC_HIP(hipcub::DeviceRadixSort::SortKeys(c.tmp.p, bytes, kin, kout, n, 0, bits, c.stream));
}
void sort_i32_pairs(Context& c, const int* kin, int* kout, const int* vin, int* vout, int n, int bits) {
    if (n <= 0) return;
    size_t bytes = 0;
    CIPC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, n, 0, bits, c.stream));
    c.tmp.reserve(bytes);
    CIPC_HIP(hipcub::DeviceRadixSort::SortPairs(c.tmp.p, bytes, kin, kout, vin, vout, n, 0, bits, c.stream));
}
void sort_u32_pairs(Context& c, const unsigned* kin, unsigned* kout, const int* vin, int* vout, int n, int bits) {
    sort_u32_pairs(kin, kout, vin, vout, n, bits, c.stream, c.tmp);
}
void sort_u32_pairs(const unsigned* kin, unsigned* kout, const int* vin, int* vout, int n, int bits, hipStream_t st,
                    DBuf<unsigned char>& tmp) {
    if (n <= 0) return;
    size_t bytes = 0;
    CIPC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, n, 0, bits, st));
    tmp.reserve(bytes);
    CIPC_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin, kout, vin, vout, n, 0, bits, st));
}
// out[i] = sum in[0..i-1]; out has n+1 entries when in has n (out[n] = total).
void exclusive_scan_int(Context& c, const int* in, int* out, int n) {
    size_t bytes = 0;
    CIPC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, c.stream));
    c.tmp.reserve(bytes);
    CIPC_HIP(hipcub::DeviceScan::ExclusiveSum(c.tmp.p, bytes, in, out, n, c.stream));
}

// --- contact block records: equal-key runs (heads from bucket_tails), their sums and nonzero masks -----
// End of the run of `key` that starts at i in the sorted keys (galloping, then bisection), so
// that the summation loop has known bounds and its loads are independent.
template <typename K>
__device__ inline int run_end(const K* __restrict__ k, int n, int i, K key) {
    int a = i, b = i + 1, st = 1;
    while (b < n && k[b] == key) { a = b; st <<= 1; b = i + st; }
    if (b > n) b = n;
    while (b - a > 1) { const int m = (a + b) >> 1; if (k[m] == key) a = m; else b = m; }
    return b;
}
// Start of every run, in run order (start[pos[i]] = i at each head), and one past the last
// keyed record after the last run, so that run o is [start[o], start[o + 1]).
__global__ void run_starts_u64(int n, const unsigned long long* __restrict__ k, const int* __restrict__ head,
                               const int* __restrict__ pos, int* __restrict__ start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i < n && head[i]) start[pos[i]] = i;
    const bool end = (i == n) ? (n == 0 || k[n - 1] != kNoKey) : (k[i] == kNoKey && (i == 0 || k[i - 1] != kNoKey));
    if (end) start[pos[n]] = i;
}
// One thread per run (unique block): the records' 3x3 values summed in record order, entries
// equal to zero skipped, and the mask of entries that received a nonzero value.
__global__ void reduce_blocks(int n, const unsigned long long* __restrict__ k, const int* __restrict__ idx,
                              const int* __restrict__ start, const int* __restrict__ pos, const double* __restrict__ val,
                              unsigned long long* __restrict__ ukey, double* __restrict__ uval, unsigned short* __restrict__ uflag) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= pos[n]) return;
    const int i = start[o], e = start[o + 1];
    const unsigned long long key = k[i];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    unsigned mask = 0;
#pragma unroll 2
    for (int j = i; j < e; ++j) {
        const double* v = val + (size_t)9 * idx[j];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const bool nz = v[c] != 0.0;
            acc[c] += nz ? v[c] : 0.0;   // exact: an accumulator starting at +0 never holds -0
            mask |= nz ? (1u << c) : 0u;
        }
    }
    ukey[o] = key;
#pragma unroll
    for (int c = 0; c < 9; ++c) uval[(size_t)9 * o + c] = acc[c];
    uflag[o] = (unsigned short)mask;
}
// --- contact gradient records, summed per vertex in record order ------------------
__global__ void add_gradient_records(int n, const int* __restrict__ k, const int* __restrict__ idx,
                                     const double* __restrict__ val, int V, double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = k[i];
    if (v < 0 || v >= V || (i > 0 && k[i - 1] == v)) return;
    const int e = run_end(k, n, i, v);
    double g0 = grad[3 * v], g1 = grad[3 * v + 1], g2 = grad[3 * v + 2];
#pragma unroll 2
    for (int j = i; j < e; ++j) {
        const double* p = val + (size_t)3 * idx[j];
        g0 += p[0]; g1 += p[1]; g2 += p[2];
    }
    grad[3 * v] = g0; grad[3 * v + 1] = g1; grad[3 * v + 2] = g2;
}
void add_gradient_records_into(Context& c, int n, const int* k, const int* idx, const double* val, double* grad) {
    if (n <= 0) return;
    hipLaunchKernelGGL(add_gradient_records, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, k, idx, val, c.sc.V, grad);
    CIPC_CHECK_LAUNCH();
}
// --- merge static and contact blocks per block column --------------------------------
__device__ inline int lower_bound_u64(const unsigned long long* a, int n, unsigned long long key) {
    int lo = 0, hi = n;
    while (lo < hi) { const int m = (lo + hi) >> 1; if (a[m] < key) lo = m + 1; else hi = m; }
    return lo;
}
__global__ void merge_count(int V, int kb, const int* __restrict__ scol, const int* __restrict__ srow,
                            const unsigned long long* __restrict__ ukey, const int* __restrict__ nu_ptr, int* __restrict__ cnt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) cnt[V] = 0;   // the scan's tail
    if (b >= V) return;
    const int nu = *nu_ptr;
    int i = scol[b], ie = scol[b + 1];
    int j = lower_bound_u64(ukey, nu, (unsigned long long)b << kb);
    const int je = lower_bound_u64(ukey, nu, (unsigned long long)(b + 1) << kb);
    int n = 0;
    while (i < ie || j < je) {
        const int ra = i < ie ? srow[i] : 0x7fffffff;
        const int rb = j < je ? (int)(ukey[j] & ((1ull << kb) - 1)) : 0x7fffffff;
        if (ra == rb) { ++i; ++j; } else if (ra < rb) ++i; else ++j;
        ++n;
    }
    cnt[b] = n;
}
__global__ void merge_fill(int V, int kb, const int* __restrict__ scol, const int* __restrict__ srow, const double* __restrict__ sbv,
                           const unsigned short* __restrict__ sbf, const unsigned long long* __restrict__ ukey,
                           const double* __restrict__ uval, const unsigned short* __restrict__ uflag, const int* __restrict__ nu_ptr,
                           const int* __restrict__ mptr, int* __restrict__ mrow, int* __restrict__ mcol, double* __restrict__ mval,
                           unsigned short* __restrict__ mflag) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= V) return;
    const int nu = *nu_ptr;
    int i = scol[b], ie = scol[b + 1];
    int j = lower_bound_u64(ukey, nu, (unsigned long long)b << kb);
    const int je = lower_bound_u64(ukey, nu, (unsigned long long)(b + 1) << kb);
    int o = mptr[b];
    while (i < ie || j < je) {
        const int ra = i < ie ? srow[i] : 0x7fffffff;
        const int rb = j < je ? (int)(ukey[j] & ((1ull << kb) - 1)) : 0x7fffffff;
        double v[9];
        unsigned f = 0;
        if (ra <= rb) {
#pragma unroll
            for (int e = 0; e < 9; ++e) v[e] = sbv[(size_t)9 * i + e];
            f = sbf[i];
        } else {
#pragma unroll
            for (int e = 0; e < 9; ++e) v[e] = 0.0;
        }
        if (rb <= ra) {
#pragma unroll
            for (int e = 0; e < 9; ++e) v[e] += uval[(size_t)9 * j + e];
            f |= uflag[j];
        }
        mrow[o] = ra < rb ? ra : rb;
        mcol[o] = b;
#pragma unroll
        for (int e = 0; e < 9; ++e) mval[(size_t)9 * o + e] = v[e];
        mflag[o] = (unsigned short)f;
        ++o;
        if (ra == rb) { ++i; ++j; } else if (ra < rb) ++i; else ++j;
    }
}
// --- scalar expansion: DOF column j = 3b + cc ------------------------------------------
// Scalar CSR.  Scalar column 3b + cc lists, for each merged block k of block column b in
// row order, the rows r of k with a nonzero flag at (r, cc) (r >= cc on the diagonal block).
// Each (block, component) pair therefore owns a contiguous run of the column; the pairs
// are laid out (b, cc, k) so one exclusive scan of their counts gives every run's start,
// and the column starts are the starts of each column's first run.
__device__ inline int pair_index(const int* __restrict__ mptr, int b, int cc, int k) {
    const int m0 = mptr[b];
    return 3 * m0 + cc * (mptr[b + 1] - m0) + (k - m0);
}
__global__ void pair_count(const int* __restrict__ nm_ptr, const int* __restrict__ mptr, const int* __restrict__ mrow,
                           const int* __restrict__ mcol, const unsigned short* __restrict__ mflag, int* __restrict__ cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nm = *nm_ptr;
    if (t > 3 * nm) return;
    if (t == 3 * nm) { cnt[t] = 0; return; }
    const int k = t / 3, cc = t % 3;
    const int b = mcol[k];
    const unsigned f = mflag[k];
    const int r0 = (mrow[k] == b) ? cc : 0;
    int n = 0;
    for (int r = r0; r < 3; ++r) n += (f >> (3 * r + cc)) & 1u;
    cnt[pair_index(mptr, b, cc, k)] = n;
}
// outer[j]: the column starts; the last one, the nonzero count, also into nnz_out
__device__ inline void column_starts(int j, int V, const int* __restrict__ mptr, const int* __restrict__ ppos,
                                     const int* __restrict__ nm_ptr, int* __restrict__ outer, int* __restrict__ nnz_out) {
    if (j > 3 * V) return;
    if (j == 3 * V) {
        const int nnz = ppos[3 * *nm_ptr];
        outer[j] = nnz;
        *nnz_out = nnz;
        return;
    }
    const int b = j / 3, cc = j % 3;
    outer[j] = ppos[pair_index(mptr, b, cc, mptr[b])];
}
__device__ inline void pair_fill(int t, const int* __restrict__ nm_ptr, const int* __restrict__ mptr, const int* __restrict__ mrow,
                                 const int* __restrict__ mcol, const double* __restrict__ mval, const unsigned short* __restrict__ mflag,
                                 const int* __restrict__ ppos, int* __restrict__ inner, double* __restrict__ val) {
    if (t >= 3 * *nm_ptr) return;
    const int k = t / 3, cc = t % 3;
    const int b = mcol[k], a = mrow[k];
    const unsigned f = mflag[k];
    int o = ppos[pair_index(mptr, b, cc, k)];
    for (int r = (a == b) ? cc : 0; r < 3; ++r) {
        if ((f >> (3 * r + cc)) & 1u) {
            inner[o] = 3 * a + r;
            val[o] = mval[(size_t)9 * k + 3 * r + cc];
            ++o;
        }
    }
}
__device__ inline void masked_rhs(int i, int V, const int* __restrict__ bt, const double* __restrict__ grad, double* __restrict__ rhs) {
    if (i >= 3 * V) return;
    rhs[i] = is_free(bt[i / 3]) ? grad[i] : 0.0;
}
// The CSR's last three steps in one launch (they read what earlier launches wrote and write
// disjoint arrays): work-groups [0, bC) the column starts, the next bP the entries of the
// merged blocks, the rest the right-hand side.
struct CsrFinish {
    int V, bC, bP;
    const int *mptr, *ppos, *nm_ptr, *mrow, *mcol, *bt;
    const double *mval, *grad;
    const unsigned short* mflag;
    int *outer, *nnz_out, *inner;
    double *val, *rhs;
};
__global__ void csr_finish(CsrFinish a) {
    int b = blockIdx.x;
    if (b < a.bC) {
        column_starts(b * blockDim.x + threadIdx.x, a.V, a.mptr, a.ppos, a.nm_ptr, a.outer, a.nnz_out);
        return;
    }
    b -= a.bC;
    if (b < a.bP) {
        pair_fill(b * blockDim.x + threadIdx.x, a.nm_ptr, a.mptr, a.mrow, a.mcol, a.mval, a.mflag, a.ppos, a.inner, a.val);
        return;
    }
    b -= a.bP;
    masked_rhs(b * blockDim.x + threadIdx.x, a.V, a.bt, a.grad, a.rhs);
}

// Collision count as compute_g_dpt returns it (ContactMechanics.cpp:2970-3032): the
// multiplicity of a point-point / point-edge entry, one for every other Self_ActiveSet
// entry.  One work-group, integer sums: the result does not depend on the order.
// out[0] = the CSR's nonzeros (nnz_dev), out[1] = the collision count: read by the host together.
// *out: compute_g_dpt's count over Self_ActiveSet (one per entry, the multiplicity for a
// point-point or point-edge map entry)
__global__ void collision_count(const int* __restrict__ n_dev, const int4* __restrict__ set, int* __restrict__ out) {
    __shared__ long long sh[256];
    long long acc = 0;
    const int n = *n_dev;
    for (int i = threadIdx.x; i < n; i += 256) {
        const int4 e = set[i];
        acc += (e.x < 0 && (e.z < 0 || e.w < 0)) ? -(long long)e.w : 1ll;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (int)sh[0];
}

__global__ void zero_ints(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}


// --- contact records bucketed by vertex: the order of a stable sort by key ----------
// The block records are bucketed by block column (the key's high field) and the gradient
// records by vertex: counts per bucket (integer atomics, order-free), one scan of both count
// arrays for the bucket starts, a scatter (atomic slot within the bucket), and one wavefront
// per bucket that orders it by (row, record) / record and writes the keys and record indices
// as a stable sort by key over all records would (records without a key last).  This
// replaces a global radix sort: every bucket is small, and the whole takes five launches.
__global__ void bucket_count(int n, const unsigned long long* __restrict__ bkey, int kb, int ng, const int* __restrict__ gkey, int V,
                             int* __restrict__ bcnt, int* __restrict__ gcnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const unsigned long long k = bkey[i];
        if (k != kNoKey) atomicAdd(bcnt + (int)(k >> kb), 1);
    }
    if (i < ng) {
        const int v = gkey[i];
        if (v >= 0 && v < V) atomicAdd(gcnt + v, 1);
    }
}
// bucket entries in one array: block records as (row, record) at [0, blocks), gradient
// records as their record index after them, where the joint scan's starts place them
__global__ void bucket_scatter(int n, const unsigned long long* __restrict__ bkey, int kb, const int* __restrict__ bstart,
                               int* __restrict__ bfill, int ng, const int* __restrict__ gkey, int V, const int* __restrict__ gstart,
                               int* __restrict__ gfill, unsigned long long* __restrict__ entries) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const unsigned long long k = bkey[i];
        if (k != kNoKey) {
            const int b = (int)(k >> kb);
            const unsigned row = (unsigned)(k & ((1ull << kb) - 1));
            entries[bstart[b] + atomicAdd(bfill + b, 1)] = ((unsigned long long)row << 32) | (unsigned)i;
        }
    }
    if (i < ng) {
        const int v = gkey[i];
        if (v >= 0 && v < V) entries[gstart[v] + atomicAdd(gfill + v, 1)] = (unsigned long long)(unsigned)i;
    }
}
constexpr int kBucketWaves = 2;
constexpr int kBucketLds = 2048;
__device__ inline unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    const unsigned lo = __shfl_xor((unsigned)v, m), hi = __shfl_xor((unsigned)(v >> 32), m);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ inline void bucket_wave_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Output of entry t of bucket bk (sorted entry e): a block record's key and index, or a
// gradient record's vertex and index.
struct BucketOut {
    int V, kb, gbase;
    unsigned long long* bkey;
    int* bidx;
    int* gkey;
    int* gidx;
    __device__ void put(int bk, int pos, unsigned long long e) const {
        if (bk < V) {
            bkey[pos] = ((unsigned long long)(unsigned)bk << kb) | (e >> 32);
            bidx[pos] = (int)(e & 0xffffffffu);
        } else {
            gkey[pos - gbase] = bk - V - 1;
            gidx[pos - gbase] = (int)(e & 0xffffffffu);
        }
    }
};
// Bucket bk = [starts[bk], starts[bk + 1]) of the 2V + 1 buckets, k > 1 entries, sorted by
// one wavefront: up to 64 entries in registers (bitonic network over the lanes), up to
// kBucketLds in LDS (bitonic network, the wave ordering its own LDS accesses between steps),
// longer buckets by one lane in place.
__device__ inline void sort_bucket(int bk, int s0, int k, int lane, unsigned long long* __restrict__ a,
                                   unsigned long long* __restrict__ keys, const BucketOut& out) {
    if (k <= 64) {
        unsigned long long v = lane < k ? keys[s0 + lane] : ~0ull;
#pragma unroll
        for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                const unsigned long long o = shfl_xor_u64(v, stride);
                const bool keep_min = ((lane & stride) == 0) == ((lane & size) == 0);
                v = keep_min ? (o < v ? o : v) : (o > v ? o : v);
            }
        if (lane < k) out.put(bk, s0 + lane, v);
        return;
    }
    if (k <= kBucketLds) {
        int P = 128;
        while (P < k) P <<= 1;
        for (int i = lane; i < P; i += 64) a[i] = i < k ? keys[s0 + i] : ~0ull;
        bucket_wave_order();
        for (int size = 2; size <= P; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = lane; t < (P >> 1); t += 64) {
                    const int i = 2 * stride * (t / stride) + (t % stride), j = i + stride;
                    const unsigned long long x = a[i], y = a[j];
                    const bool asc = (i & size) == 0;
                    if ((x > y) == asc) { a[i] = y; a[j] = x; }
                }
                bucket_wave_order();
            }
        for (int i = lane; i < k; i += 64) out.put(bk, s0 + i, a[i]);
        bucket_wave_order();   // the LDS is the next bucket's
        return;
    }
    if (lane == 0) {
        for (int i = s0 + 1; i < s0 + k; ++i) {
            const unsigned long long x = keys[i];
            int j = i - 1;
            while (j >= s0 && keys[j] > x) { keys[j + 1] = keys[j]; --j; }
            keys[j + 1] = x;
        }
        for (int i = s0; i < s0 + k; ++i) out.put(bk, i, keys[i]);
    }
}
// Each wavefront takes bpw consecutive buckets (1 or 64): a lane writes its bucket's entry
// when it has one, and the wave sorts the longer ones together, one after the other.  64
// when the record slots are fewer than the buckets (most buckets then have none or one: a
// batch of scenes with few contacts), else 1, so that long buckets (vertices in many
// contacts, often neighbours) are not sorted one after the other by one wavefront.
__global__ __launch_bounds__(64 * kBucketWaves) void bucket_sort(int nbk, int bpw, const int* __restrict__ starts,
                                                                 unsigned long long* __restrict__ keys, BucketOut out) {
    __shared__ unsigned long long sh[kBucketWaves][kBucketLds];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int bk0 = (blockIdx.x * kBucketWaves + w) * bpw;
    if (bk0 >= nbk) return;   // the whole wavefront: no block-wide barrier follows
    out.gbase = starts[out.V + 1];   // the gradient entries follow the block entries
    const int bk = bk0 + lane;
    int s0 = 0, k = 0;
    if (lane < bpw && bk < nbk) {
        s0 = starts[bk];
        k = starts[bk + 1] - s0;
    }
    if (k == 1) out.put(bk, s0, keys[s0]);
    unsigned long long big = __ballot(k > 1);
    while (big) {
        const int l = __ffsll((long long)big) - 1;
        big &= big - 1;
        sort_bucket(bk0 + l, __shfl(s0, l), __shfl(k, l), lane, sh[w], keys, out);
    }
}
// Records without a key, after the keyed ones: block keys kNoKey, gradient keys past every vertex.
// After the bucket sort, in one launch: the unkeyed records after the keyed ones (as a stable
// sort leaves them), the run heads of the sorted block keys (head[nb] = 0: the scan's total
// is the run count), and the bucket counters cleared for the next assembly (nothing after
// the bucket scatter reads them).  Keys past the keyed ones are the sentinel whether or not
// this launch has written them yet.
__global__ void bucket_tails(int nb, int ng, const int* __restrict__ starts, int V, unsigned long long* __restrict__ bkey,
                             int* __restrict__ bidx, int* __restrict__ gkey, int* __restrict__ gidx, int* __restrict__ head,
                             int* __restrict__ counters, int ncounters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int tb = starts[V], tg = starts[2 * V + 1] - starts[V + 1];
    if (i < nb) {
        const unsigned long long k = i < tb ? bkey[i] : kNoKey;
        if (i >= tb) { bkey[i] = kNoKey; bidx[i] = i; }
        const unsigned long long kp = i == 0 ? 0ull : (i - 1 < tb ? bkey[i - 1] : kNoKey);
        head[i] = (k != kNoKey && (i == 0 || k != kp)) ? 1 : 0;
    } else if (i == nb) {
        head[nb] = 0;
    }
    if (i >= tg && i < ng) { gkey[i] = 0x7fffffff; gidx[i] = i; }
    if (i < ncounters) counters[i] = 0;
}

// The nonzero and collision counts (c.coll) to pinned memory, for the host's next
// synchronisation.
void publish_nnz(Context& c, hipStream_t st) {
    c.pin_nnz.reserve(2);
    CIPC_HIP(hipMemcpyAsync(c.pin_nnz.p, c.coll.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    c.nnz_pending = true;
}
// The collision count of the assembly's set (c.coll[1]), on stream st.
void collision_count_launch(Context& c, hipStream_t st) {
    c.coll.resize(2);
    hipLaunchKernelGGL(collision_count, dim3(1), dim3(256), 0, st, c.dcnt.p + kDcSelf, c.self_set.p, c.coll.p + 1);
    CIPC_CHECK_LAUNCH();
}

void finalize_csr(Context& c, bool count_collisions) {
    Scene& s = c.sc;
    const int V = s.V;
    // contact block records
    const int nb = (int)c.ckey.n;
    c.ckey_sorted.resize(std::max(nb, 1)); c.cidx_sorted.resize(std::max(nb, 1));
    c.iscratch2.resize(std::max<size_t>((size_t)3 * V + 1, (size_t)nb + 2));
    c.counters.resize(4);
    c.ukey.resize(std::max(nb, 1)); c.uval.resize((size_t)9 * std::max(nb, 1)); c.uflag.resize(std::max(nb, 1));
    int* head = c.iscratch2.p;
    int* pos = nullptr;
    c.iscratch.resize((size_t)std::max(nb, 1) + 1);
    pos = c.iscratch.p;
    const int ng = (int)c.gkey.n;
    // the records of both kinds in key order, bucketed by vertex (counts of the block columns
    // and of the gradient vertices adjacent: one scan)
    {
        c.bucket_cnt.resize(4 * ((size_t)V + 1));
        int* bcnt = c.bucket_cnt.p;
        int* gcnt = bcnt + (V + 1);
        int* bfill = gcnt + (V + 1);
        int* gfill = bfill + (V + 1);
        c.bucket_start.resize(2 * ((size_t)V + 1));
        int* bstart = c.bucket_start.p;
        int* gstart = bstart + (V + 1);
        c.bucket_entries.resize((size_t)nb + ng + 1);
        c.gkey_sorted.resize(std::max(ng, 1)); c.gidx_sorted.resize(std::max(ng, 1));
        // the counters are left cleared by the previous assembly's bucket_tails; cleared here
        // when that is not known (a new buffer, a larger scene, an enqueue cut short)
        const size_t ncnt = 4 * ((size_t)V + 1);
        if (c.bucket_cnt_clear < ncnt || c.bucket_cnt_ptr != bcnt)
            hipLaunchKernelGGL(zero_ints, dim3(blocks_for(ncnt, 256)), dim3(256), 0, c.stream, bcnt, (int)ncnt);
        c.bucket_cnt_clear = 0;
        c.bucket_cnt_ptr = bcnt;
        const int nmax = std::max(nb, ng);
        if (nmax > 0)
            hipLaunchKernelGGL(bucket_count, dim3(blocks_for(nmax, 256)), dim3(256), 0, c.stream, nb, c.ckey.p, s.kb, ng, c.gkey.p, V,
                               bcnt, gcnt);
        CIPC_CHECK_LAUNCH();
        exclusive_scan_int(c, bcnt, bstart, 2 * (V + 1));
        if (nmax > 0) {
            hipLaunchKernelGGL(bucket_scatter, dim3(blocks_for(nmax, 256)), dim3(256), 0, c.stream, nb, c.ckey.p, s.kb, bstart, bfill,
                               ng, c.gkey.p, V, gstart, gfill, c.bucket_entries.p);
            const BucketOut bo{V, s.kb, 0, c.ckey_sorted.p, c.cidx_sorted.p, c.gkey_sorted.p, c.gidx_sorted.p};
            const int bpw = (size_t)nb + ng < 2 * (size_t)V + 1 ? 64 : 1;
            hipLaunchKernelGGL(bucket_sort, dim3(blocks_for(2 * V + 1, kBucketWaves * bpw)), dim3(64 * kBucketWaves), 0,
                               c.stream, 2 * V + 1, bpw, bstart, c.bucket_entries.p, bo);
            const size_t nt = std::max<size_t>(std::max<size_t>((size_t)nb + 1, (size_t)ng), ncnt);
            hipLaunchKernelGGL(bucket_tails, dim3(blocks_for(nt, 256)), dim3(256), 0, c.stream, nb, ng, bstart, V, c.ckey_sorted.p,
                               c.cidx_sorted.p, c.gkey_sorted.p, c.gidx_sorted.p, head, bcnt, (int)ncnt);
        }
        CIPC_CHECK_LAUNCH();
        c.bucket_cnt_clear = ncnt;   // by bucket_tails, or never touched (no records)
    }
    if (nb > 0) {
        exclusive_scan_int(c, head, pos, nb + 1);
        c.run_start.resize((size_t)nb + 2);
        hipLaunchKernelGGL(run_starts_u64, dim3(blocks_for(nb + 1, 256)), dim3(256), 0, c.stream, nb, c.ckey_sorted.p, head, pos,
                           c.run_start.p);
        hipLaunchKernelGGL(reduce_blocks, dim3(blocks_for(nb, 256)), dim3(256), 0, c.stream, nb, c.ckey_sorted.p,
                           c.cidx_sorted.p, c.run_start.p, pos, c.cval.p, c.ukey.p, c.uval.p, c.uflag.p);
        CIPC_CHECK_LAUNCH();
    } else {
        hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, c.stream, c.counters.p, 1);
        CIPC_CHECK_LAUNCH();
    }
    // the number of unique contact blocks: the run scan's total (zero without records)
    const int* nu_ptr = nb > 0 ? pos + nb : c.counters.p;
    // contact gradient records
    if (ng > 0) {
        hipLaunchKernelGGL(add_gradient_records, dim3(blocks_for(ng, 256)), dim3(256), 0, c.stream, ng, c.gkey_sorted.p,
                           c.gidx_sorted.p, c.gval.p, V, c.grad.p);
        CIPC_CHECK_LAUNCH();
    }
    // merged block pattern, sized by its bound (static blocks + unique contact blocks <= records)
    c.mptr.resize((size_t)V + 1);
    DBuf<int>& cnt = c.iscratch2;
    hipLaunchKernelGGL(merge_count, dim3(blocks_for(V, 256)), dim3(256), 0, c.stream, V, s.kb, c.sp.col_ptr.p, c.sp.row.p, c.ukey.p,
                       nu_ptr, cnt.p);
    CIPC_CHECK_LAUNCH();
    exclusive_scan_int(c, cnt.p, c.mptr.p, V + 1);
    const size_t mcap = (size_t)c.sp.nblocks + (size_t)nb + 1;
    c.mrow.resize(mcap); c.mcol.resize(mcap); c.mval.resize(9 * mcap); c.mflag.resize(mcap);
    hipLaunchKernelGGL(merge_fill, dim3(blocks_for(V, 256)), dim3(256), 0, c.stream, V, s.kb, c.sp.col_ptr.p, c.sp.row.p, c.sbv.p,
                       c.sbf.p, c.ukey.p, c.uval.p, c.uflag.p, nu_ptr, c.mptr.p, c.mrow.p, c.mcol.p, c.mval.p, c.mflag.p);
    CIPC_CHECK_LAUNCH();
    // scalar CSR: (block, component) counts -> one scan -> column starts and runs
    const int* nm_ptr = c.mptr.p + V;
    c.pcnt.resize(3 * mcap + 1); c.ppos.resize(3 * mcap + 1);
    hipLaunchKernelGGL(pair_count, dim3(blocks_for(3 * mcap + 1, 256)), dim3(256), 0, c.stream, nm_ptr, c.mptr.p, c.mrow.p, c.mcol.p,
                       c.mflag.p, c.pcnt.p);
    CIPC_CHECK_LAUNCH();
    // the scan covers the bound; entries past 3 nm + 1 are never read
    exclusive_scan_int(c, c.pcnt.p, c.ppos.p, (int)(3 * mcap + 1));
    c.outer.resize((size_t)3 * V + 1);
    c.coll.resize(2);
    c.inner.resize(9 * mcap); c.val.resize(9 * mcap);
    c.rhs.resize((size_t)3 * V);
    {
        CsrFinish a;
        a.V = V; a.bC = (int)blocks_for(3 * V + 1, 256); a.bP = (int)blocks_for(3 * mcap, 256);
        a.mptr = c.mptr.p; a.ppos = c.ppos.p; a.nm_ptr = nm_ptr; a.mrow = c.mrow.p; a.mcol = c.mcol.p; a.bt = s.bt.p;
        a.mval = c.mval.p; a.grad = c.grad.p; a.mflag = c.mflag.p; a.outer = c.outer.p; a.nnz_out = c.coll.p;
        a.inner = c.inner.p; a.val = c.val.p; a.rhs = c.rhs.p;
        hipLaunchKernelGGL(csr_finish, dim3(a.bC + a.bP + (int)blocks_for(3 * V, 256)), dim3(256), 0, c.stream, a);
        CIPC_CHECK_LAUNCH();
    }
    // the counts are read with the caller's next synchronisation (the collision count is
    // launched here, or by the caller on a stream of its own before it publishes them); the
    // entries are sized by their bound (9 per merged block)
    if (count_collisions) {
        collision_count_launch(c, c.stream);
        publish_nnz(c, c.stream);
    }
    CIPC_CHECK_LAUNCH();
}

}  // namespace cipc
