// lsgpu_sah.hip -- SAH rebuild of the LBVH's inner nodes on the GPU.
//
// The same tree as the host rebuild (lsgpu_sah.cpp, lsg_bvh_rebuild_sah), built breadth
// first: the leaves of the LBVH (triangle ranges of at most leaf_size, in the order of their
// triangles) are rebuilt top-down, one wave per node of a level.  A wave bins its leaves'
// centroids (32 bins per axis, as many as leaves below 32) into LDS, one lane picks the
// cheapest split as the host does (area(left)·n(left) + area(right)·n(right) over the three
// axes; equal centroids split the range in half), and the wave partitions its leaves stably
// into the other buffer, collecting the children's boxes and centroid bounds.  Box and bound
// reductions are exact (min/max over floats as order-preserving integers), the split
// arithmetic is the host's, and nodes are numbered in preorder (left child = node + 1, right
// child = node + leaves on the left), so the result equals the host's node for node.
// Tasks of 2^15 leaves or more (the top levels) take the same steps across the chip:
// many workgroups bin, one lane splits, a scan of left/right flags partitions.
#include "lsgpu_internal.h"
#include <vector>

namespace lsg {
namespace {

constexpr int kSahBins = 32;

struct SahLeaf { int32_t ref; float b[6]; float c[3]; };
struct SahTask { uint32_t begin, end, id, depth; float clo[3], chi[3]; };

// floats as integers of the same order (-0 as +0), for atomic min/max
__device__ inline int f2o(float x)
{
    int i = __float_as_int(x);
    if (i == (int)0x80000000) i = 0;
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ inline float o2f(int i) { return __int_as_float(i >= 0 ? i : (i ^ 0x7fffffff)); }

__device__ inline bool reachable(int i, const int* first, const int* last, uint32_t leafmax)
{
    return (uint32_t)(last[i] - first[i] + 1) > leafmax;   // the root has every triangle
}

// leaves of the LBVH: each reachable node's children of at most leafmax triangles
__global__ __launch_bounds__(256) void k_sah_mark(int ni, const int* __restrict__ first, const int* __restrict__ last,
                                                  const int* __restrict__ split, uint32_t leafmax, uint32_t* __restrict__ flag)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ni || !reachable(i, first, last, leafmax)) return;
    const int g = split[i];
    if ((uint32_t)(g - first[i] + 1) <= leafmax) flag[first[i]] = 1u;
    if ((uint32_t)(last[i] - g) <= leafmax) flag[g + 1] = 1u;
}

__global__ __launch_bounds__(256) void k_sah_gather(int ni, const int* __restrict__ first, const int* __restrict__ last,
                                                    const int* __restrict__ split, uint32_t leafmax, const Node* __restrict__ nodes,
                                                    const uint32_t* __restrict__ pos, SahLeaf* __restrict__ L)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ni || !reachable(i, first, last, leafmax)) return;
    const int g = split[i];
    const Node q = nodes[i];
    const float f[16] = {q.q0.x, q.q0.y, q.q0.z, q.q0.w, q.q1.x, q.q1.y, q.q1.z, q.q1.w,
                         q.q2.x, q.q2.y, q.q2.z, q.q2.w, q.q3.x, q.q3.y, q.q3.z, q.q3.w};
    for (int c = 0; c < 2; c++) {
        const int cf = c == 0 ? first[i] : g + 1;
        const int cl = c == 0 ? g : last[i];
        if ((uint32_t)(cl - cf + 1) > leafmax) continue;
        SahLeaf lf;
        lf.ref = __float_as_int(f[12 + c]);
        for (int k = 0; k < 6; k++) lf.b[k] = f[6 * c + k];
        for (int a = 0; a < 3; a++) lf.c[a] = (lf.b[2 * a] + lf.b[2 * a + 1]) * 0.5f;
        L[pos[cf]] = lf;
    }
}

// centroid bounds of all leaves (the root task's), as ordered integers: lo x y z, hi x y z
__global__ __launch_bounds__(256) void k_sah_root(const SahLeaf* __restrict__ L, uint32_t n, int* __restrict__ cb)
{
    __shared__ int s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? f2o(3.0e38f) : f2o(-3.0e38f);
    __syncthreads();
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        for (int a = 0; a < 3; a++) {
            atomicMin(&s[a], f2o(L[k].c[a]));
            atomicMax(&s[3 + a], f2o(L[k].c[a]));
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&cb[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&cb[threadIdx.x], s[threadIdx.x]);
}

__global__ void k_sah_root_task(const int* __restrict__ cb, uint32_t n, SahTask* __restrict__ t)
{
    if (threadIdx.x) return;
    SahTask r;
    r.begin = 0; r.end = n; r.id = 0; r.depth = 1;
    for (int a = 0; a < 3; a++) { r.clo[a] = o2f(cb[a]); r.chi[a] = o2f(cb[3 + a]); }
    *t = r;
}

// bins per axis of a task and the binning scale (0 on an axis whose centroids coincide)
__device__ inline int task_bins(const SahTask& T, bool* ok, float* sc)
{
    const uint32_t m = T.end - T.begin;
    const int nb = (int)min((uint32_t)kSahBins, max(4u, m));
    for (int a = 0; a < 3; a++) {
        ok[a] = T.chi[a] > T.clo[a];
        sc[a] = ok[a] ? (float)nb / (T.chi[a] - T.clo[a]) : 0.f;
    }
    return nb;
}

__device__ inline int bin_of(const SahLeaf& f, const SahTask& T, int a, float sc, int nb)
{
    return min((int)((f.c[a] - T.clo[a]) * sc), nb - 1);
}

// the host's split search (lsgpu_sah.cpp) on a task's bins: dec = (axis or -1 for halves,
// last bin on the left, leaves on the left)
__device__ inline void sah_choose(const int (*cnt)[kSahBins], const int (*box)[kSahBins][6], int nb, const bool* ok, uint32_t m,
                                  int* dec)
{
    float best = 0.f;
    int bax = -1, bk = -1;
    for (int a = 0; a < 3; a++) {
        if (!ok[a]) continue;
        float ra[kSahBins];
        uint32_t rc[kSahBins];
        float acc[6] = {3.0e38f, -3.0e38f, 3.0e38f, -3.0e38f, 3.0e38f, -3.0e38f};
        uint32_t c = 0;
        for (int k = nb - 1; k >= 0; k--) {
            for (int e = 0; e < 6; e++) {
                const float v = o2f(box[a][k][e]);
                acc[e] = (e & 1) ? fmaxf(acc[e], v) : fminf(acc[e], v);
            }
            c += (uint32_t)cnt[a][k];
            const float dx = acc[1] - acc[0], dy = acc[3] - acc[2], dz = acc[5] - acc[4];
            ra[k] = c ? dx * dy + dy * dz + dz * dx : 0.f;
            rc[k] = c;
        }
        for (int e = 0; e < 6; e++) acc[e] = (e & 1) ? -3.0e38f : 3.0e38f;
        c = 0;
        for (int k = 0; k < nb - 1; k++) {
            for (int e = 0; e < 6; e++) {
                const float v = o2f(box[a][k][e]);
                acc[e] = (e & 1) ? fmaxf(acc[e], v) : fminf(acc[e], v);
            }
            c += (uint32_t)cnt[a][k];
            if (!c || !rc[k + 1]) continue;
            const float dx = acc[1] - acc[0], dy = acc[3] - acc[2], dz = acc[5] - acc[4];
            const float cost = (dx * dy + dy * dz + dz * dx) * (float)c + ra[k + 1] * (float)rc[k + 1];
            if (bax < 0 || cost < best) { best = cost; bax = a; bk = k; }
        }
    }
    uint32_t nl = m / 2;
    if (bax >= 0) {
        nl = 0;
        for (int k = 0; k <= bk; k++) nl += (uint32_t)cnt[bax][k];
    }
    dec[0] = bax; dec[1] = bk; dec[2] = (int)nl;
}

// a split task's node (children's boxes, a leaf child's ref from dst) and its children of
// two or more leaves as the next level's tasks
__device__ inline void sah_emit(const SahTask& T, uint32_t nl, const int (*child)[12], const SahLeaf* dst, Node* out, SahTask* next,
                                uint32_t* nnext, uint32_t* depth)
{
    atomicMax(depth, T.depth);
    float nodef[16];
    // the children of two or more leaves become tasks, appended with one atomic
    const uint32_t nkids = (nl >= 2 ? 1u : 0u) + (T.end - T.begin - nl >= 2 ? 1u : 0u);
    uint32_t slot = nkids ? atomicAdd(nnext, nkids) : 0u;
    for (int s = 0; s < 2; s++) {
        const uint32_t b = s == 0 ? T.begin : T.begin + nl, e = s == 0 ? T.begin + nl : T.end;
        const uint32_t ms = e - b;
        const uint32_t id = s == 0 ? T.id + 1 : T.id + nl;
        const int32_t ref = ms == 1 ? dst[b].ref : (int32_t)id;
        for (int k = 0; k < 6; k++) nodef[6 * s + k] = o2f(child[s][k]);
        nodef[12 + s] = __int_as_float(ref);
        if (ms >= 2) {
            SahTask c;
            c.begin = b; c.end = e; c.id = id; c.depth = T.depth + 1;
            for (int a = 0; a < 3; a++) { c.clo[a] = o2f(child[s][6 + a]); c.chi[a] = o2f(child[s][9 + a]); }
            next[slot++] = c;
        }
    }
    nodef[14] = 0.f; nodef[15] = 0.f;
    Node nd;
    nd.q0 = make_float4(nodef[0], nodef[1], nodef[2], nodef[3]);
    nd.q1 = make_float4(nodef[4], nodef[5], nodef[6], nodef[7]);
    nd.q2 = make_float4(nodef[8], nodef[9], nodef[10], nodef[11]);
    nd.q3 = make_float4(nodef[12], nodef[13], nodef[14], nodef[15]);
    out[T.id] = nd;
}

__device__ inline void bins_init(int* cnt, int* box, uint32_t tid, uint32_t nthreads)
{
    const int kLo = f2o(3.0e38f), kHi = f2o(-3.0e38f);
    for (uint32_t k = tid; k < 3 * kSahBins; k += nthreads) {
        cnt[k] = 0;
        for (int e = 0; e < 6; e++) box[6 * k + e] = (e & 1) ? kHi : kLo;
    }
}

__device__ inline void child_init(int* child, uint32_t tid)
{
    // per child: box lo/hi x y z (6), centroid lo x y z, hi x y z (6)
    if (tid < 24) child[tid] = ((tid % 12) < 6 ? ((tid & 1) ? f2o(-3.0e38f) : f2o(3.0e38f)) : ((tid % 12) < 9 ? f2o(3.0e38f) : f2o(-3.0e38f)));
}

__device__ inline void bin_leaf(const SahLeaf& f, const SahTask& T, const bool* ok, const float* sc, int nb, int (*cnt)[kSahBins],
                                int (*box)[kSahBins][6])
{
    for (int a = 0; a < 3; a++) {
        if (!ok[a]) continue;
        const int j = bin_of(f, T, a, sc[a], nb);
        atomicAdd(&cnt[a][j], 1);
        for (int e = 0; e < 6; e++) {
            if (e & 1) atomicMax(&box[a][j][e], f2o(f.b[e]));
            else atomicMin(&box[a][j][e], f2o(f.b[e]));
        }
    }
}

__device__ inline void child_add(int* cbx, const SahLeaf& f)
{
    for (int e = 0; e < 6; e++) {
        if (e & 1) atomicMax(&cbx[e], f2o(f.b[e]));
        else atomicMin(&cbx[e], f2o(f.b[e]));
    }
    for (int a = 0; a < 3; a++) {
        atomicMin(&cbx[6 + a], f2o(f.c[a]));
        atomicMax(&cbx[9 + a], f2o(f.c[a]));
    }
}

// one node of a level per wave: bins, split, stable partition src -> dst, children
__global__ __launch_bounds__(64) void k_sah_level(const SahTask* __restrict__ tasks, const SahLeaf* __restrict__ src,
                                                  SahLeaf* __restrict__ dst, Node* __restrict__ out, SahTask* __restrict__ next,
                                                  uint32_t* __restrict__ nnext, uint32_t* __restrict__ depth)
{
    __shared__ int s_cnt[3][kSahBins];
    __shared__ int s_box[3][kSahBins][6];
    __shared__ int s_child[2][12];
    __shared__ int s_dec[3];                // axis (-1: halves), bin, leaves on the left
    const uint32_t lane = threadIdx.x;
    const SahTask T = tasks[blockIdx.x];
    const uint32_t m = T.end - T.begin;
    bool ok[3];
    float sc[3];
    const int nb = task_bins(T, ok, sc);
    bins_init(&s_cnt[0][0], &s_box[0][0][0], lane, 64);
    child_init(&s_child[0][0], lane);
    __syncthreads();
    for (uint32_t k = T.begin + lane; k < T.end; k += 64) bin_leaf(src[k], T, ok, sc, nb, s_cnt, s_box);
    __syncthreads();
    if (lane == 0) sah_choose(s_cnt, s_box, nb, ok, m, s_dec);
    __syncthreads();
    const int bax = s_dec[0], bk = s_dec[1];
    const uint32_t nl = (uint32_t)s_dec[2];
    uint32_t ldone = 0, rdone = 0;
    for (uint32_t base = T.begin; base < T.end; base += 64) {
        const uint32_t k = base + lane;
        const bool valid = k < T.end;
        SahLeaf f;
        bool left = false;
        if (valid) {
            f = src[k];
            left = bax < 0 ? (k - T.begin < m / 2) : (bin_of(f, T, bax, sc[bax], nb) <= bk);
        }
        const uint64_t lm = __ballot(valid && left), rm = __ballot(valid && !left);
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
        if (valid) {
            const uint32_t at = left ? T.begin + ldone + (uint32_t)__popcll(lm & below)
                                     : T.begin + nl + rdone + (uint32_t)__popcll(rm & below);
            dst[at] = f;
            child_add(s_child[left ? 0 : 1], f);
        }
        ldone += (uint32_t)__popcll(lm);
        rdone += (uint32_t)__popcll(rm);
    }
    __syncthreads();
    if (lane == 0) sah_emit(T, nl, s_child, dst, out, next, nnext, depth);
}

// ---- one big task across the chip ------------------------------------------------------
// The top levels have a handful of tasks over millions of leaves: each is binned by many
// workgroups (LDS bins merged into global ones with the same exact integer min/max and
// counts), split by one lane with the same search, partitioned stably by a scan of its
// left/right flags, and emitted as k_sah_level emits.
constexpr uint32_t kBigTask = 1u << 15;

struct SahBig {
    int cnt[3][kSahBins];
    int box[3][kSahBins][6];
    int child[2][12];
    int dec[3];
};

__global__ __launch_bounds__(256) void k_big_init(SahBig* __restrict__ B)
{
    bins_init(&B->cnt[0][0], &B->box[0][0][0], threadIdx.x, 256);
    child_init(&B->child[0][0], threadIdx.x);
}

__global__ __launch_bounds__(256) void k_big_bin(const SahTask* __restrict__ task, const SahLeaf* __restrict__ src, SahBig* __restrict__ B)
{
    __shared__ int s_cnt[3][kSahBins];
    __shared__ int s_box[3][kSahBins][6];
    const SahTask T = *task;
    bool ok[3];
    float sc[3];
    const int nb = task_bins(T, ok, sc);
    bins_init(&s_cnt[0][0], &s_box[0][0][0], threadIdx.x, 256);
    __syncthreads();
    for (uint32_t k = T.begin + blockIdx.x * 256 + threadIdx.x; k < T.end; k += gridDim.x * 256) bin_leaf(src[k], T, ok, sc, nb, s_cnt, s_box);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 3 * kSahBins; k += 256) {
        const int c = (&s_cnt[0][0])[k];
        if (!c) continue;
        atomicAdd(&B->cnt[0][0] + k, c);
        for (int e = 0; e < 6; e++) {
            int* g = &B->box[0][0][0] + 6 * k + e;
            const int v = (&s_box[0][0][0])[6 * k + e];
            if (e & 1) atomicMax(g, v); else atomicMin(g, v);
        }
    }
}

__global__ void k_big_split(const SahTask* __restrict__ task, SahBig* __restrict__ B)
{
    if (threadIdx.x) return;
    const SahTask T = *task;
    bool ok[3];
    float sc[3];
    const int nb = task_bins(T, ok, sc);
    sah_choose(B->cnt, B->box, nb, ok, T.end - T.begin, B->dec);
}

__global__ __launch_bounds__(256) void k_big_flag(const SahTask* __restrict__ task, const SahLeaf* __restrict__ src,
                                                  const SahBig* __restrict__ B, uint32_t* __restrict__ flag)
{
    const SahTask T = *task;
    const uint32_t m = T.end - T.begin;
    bool ok[3];
    float sc[3];
    const int nb = task_bins(T, ok, sc);
    const int bax = B->dec[0], bk = B->dec[1];
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256)
        flag[i] = (bax < 0 ? i < m / 2 : bin_of(src[T.begin + i], T, bax, sc[bax], nb) <= bk) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_big_scatter(const SahTask* __restrict__ task, const SahLeaf* __restrict__ src,
                                                     SahLeaf* __restrict__ dst, const uint32_t* __restrict__ flag,
                                                     const uint32_t* __restrict__ prefix, SahBig* __restrict__ B)
{
    __shared__ int s_child[2][12];
    const SahTask T = *task;
    const uint32_t m = T.end - T.begin, nl = (uint32_t)B->dec[2];
    child_init(&s_child[0][0], threadIdx.x);
    __syncthreads();
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
        const SahLeaf f = src[T.begin + i];
        const bool left = flag[i] != 0;
        dst[left ? T.begin + prefix[i] : T.begin + nl + (i - prefix[i])] = f;
        child_add(s_child[left ? 0 : 1], f);
    }
    __syncthreads();
    if (threadIdx.x < 24) {
        const uint32_t k = threadIdx.x, e = k % 12;
        int* g = &B->child[0][0] + k;
        const int v = (&s_child[0][0])[k];
        const bool hi = e < 6 ? (e & 1) != 0 : e >= 9;
        if (hi) atomicMax(g, v); else atomicMin(g, v);
    }
}

__global__ void k_big_emit(const SahTask* __restrict__ task, const SahLeaf* __restrict__ dst, const SahBig* __restrict__ B,
                           Node* __restrict__ out, SahTask* __restrict__ next, uint32_t* __restrict__ nnext, uint32_t* __restrict__ depth)
{
    if (threadIdx.x) return;
    sah_emit(*task, (uint32_t)B->dec[2], B->child, dst, out, next, nnext, depth);
}

// ---- BVH2 -> BVH4 collapse ------------------------------------------------------------

struct C4Work { int32_t i2; uint32_t slot, P; };

// heights (inner nodes on the longest path down, a node over two leaves = 1), one level
// per launch: launch k finishes the nodes whose inner children were finished by earlier
// launches (H in 1..k-1), so H = k is the height, and every box of H read was written
// before the launch began (no fences)
__global__ __launch_bounds__(256) void k_c4_height(uint32_t nn, const Node* __restrict__ nodes, uint32_t* __restrict__ H, uint32_t k)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn || H[i]) return;
    const int32_t r0 = __float_as_int(nodes[i].q3.x), r1 = __float_as_int(nodes[i].q3.y);
    if (r0 >= 0) { const uint32_t h = H[r0]; if (h == 0 || h >= k) return; }
    if (r1 >= 0) { const uint32_t h = H[r1]; if (h == 0 || h >= k) return; }
    H[i] = k;
}

__device__ inline float box_area(const float* b)
{
    const float dx = b[1] - b[0], dy = b[3] - b[2], dz = b[5] - b[4];
    return dx * dy + dy * dz + dz * dx;
}

__device__ inline void kids2(const Node& q, int32_t* ref, float (*b)[6])
{
    ref[0] = __float_as_int(q.q3.x); ref[1] = __float_as_int(q.q3.y);
    b[0][0] = q.q0.x; b[0][1] = q.q0.y; b[0][2] = q.q0.z; b[0][3] = q.q0.w; b[0][4] = q.q1.x; b[0][5] = q.q1.y;
    b[1][0] = q.q1.z; b[1][1] = q.q1.w; b[1][2] = q.q2.x; b[1][3] = q.q2.y; b[1][4] = q.q2.z; b[1][5] = q.q2.w;
}

// one BVH4 node per work item: the BVH2 node's children, the inner child of largest area
// opened while fewer than four and while the stack budget allows one more push (the pushes
// above plus the children minus one, plus the tallest subtree below, within the budget)
__global__ __launch_bounds__(256) void k_c4_open(const C4Work* __restrict__ work, uint32_t nw, const Node* __restrict__ nodes,
                                                 const uint32_t* __restrict__ H, uint32_t budget, Node4* __restrict__ out4,
                                                 uint32_t* __restrict__ cnt, uint32_t* __restrict__ pnext, uint32_t* __restrict__ need)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nw) return;
    const C4Work w = work[i];
    int32_t ref[4];
    float b[4][6];
    kids2(nodes[w.i2], ref, b);
    int m = 2;
    while (m < 4) {
        int best = -1; float ba = -1.f;
        for (int k = 0; k < m; k++) if (ref[k] >= 0 && box_area(b[k]) > ba) { ba = box_area(b[k]); best = k; }
        if (best < 0) break;
        int32_t gr[2]; float gb[2][6];
        kids2(nodes[ref[best]], gr, gb);
        uint32_t hmax = 0;
        for (int k = 0; k < m; k++) if (k != best && ref[k] >= 0) hmax = max(hmax, H[ref[k]]);
        for (int j = 0; j < 2; j++) if (gr[j] >= 0) hmax = max(hmax, H[gr[j]]);
        if (w.P + (uint32_t)m + hmax > budget) break;
        ref[best] = gr[0]; for (int e = 0; e < 6; e++) b[best][e] = gb[0][e];
        ref[m] = gr[1]; for (int e = 0; e < 6; e++) b[m][e] = gb[1][e];
        m++;
    }
    const uint32_t P = w.P + (uint32_t)m - 1;
    atomicMax(need, P);
    Node4 nd;
    float* f = (float*)nd.q;
    int32_t* r = (int32_t*)&nd.q[6];
    uint32_t inner = 0;
    for (int k = 0; k < 4; k++) {
        for (int a = 0; a < 6; a++) f[4 * a + k] = k < m ? b[k][a] : 0.f;
        r[k] = k < m ? ref[k] : kNoChild;          // inner children: their BVH2 node until linked
        if (k < m && ref[k] >= 0) inner++;
    }
    nd.q[7] = make_float4(0.f, 0.f, 0.f, 0.f);
    out4[w.slot] = nd;
    cnt[i] = inner;
    pnext[i] = P;
}

// inner children get the next slots in work order (base + the item's prefix), and become
// the next level's work
__global__ __launch_bounds__(256) void k_c4_link(const C4Work* __restrict__ work, uint32_t nw, const uint32_t* __restrict__ prefix,
                                                 const uint32_t* __restrict__ pnext, uint32_t base, Node4* __restrict__ out4,
                                                 C4Work* __restrict__ next)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nw) return;
    const C4Work w = work[i];
    int32_t* r = (int32_t*)&out4[w.slot].q[6];
    uint32_t j = prefix[i];
    for (int k = 0; k < 4; k++) {
        const int32_t c = r[k];
        if (c < 0 || c == kNoChild) continue;
        const uint32_t s = base + j;
        next[j] = C4Work{c, s, pnext[i]};
        r[k] = (int32_t)s;
        j++;
    }
}

inline void sfree(void* p) { if (p) (void)hipFree(p); }

} // namespace

int sah_rebuild_device(const int* d_first, const int* d_last, const int* d_split, int nprim, uint32_t leafmax, const Node* d_nodes,
                       Node** out_nodes, uint32_t* out_count, uint32_t* out_depth)
{
    *out_nodes = nullptr; *out_count = 0; *out_depth = 0;
    const int ni = nprim - 1;
    if (ni < 1) return 1;
    uint32_t *flag = nullptr, *pos = nullptr, *tmp = nullptr, *total = nullptr, *cnt = nullptr, *dep = nullptr;
    int* cb = nullptr;
    SahLeaf *L0 = nullptr, *L1 = nullptr;
    SahTask *T0 = nullptr, *T1 = nullptr, *tsmall = nullptr;
    SahBig* big = nullptr;
    Node* out = nullptr;
    auto cleanup = [&]() {
        sfree(flag); sfree(pos); sfree(tmp); sfree(total); sfree(cnt); sfree(dep); sfree(cb);
        sfree(L0); sfree(L1); sfree(T0); sfree(T1); sfree(tsmall); sfree(big);
    };
#define SK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); sfree(out); return hip_fail(e_, #x, __LINE__); } } while (0)
    const unsigned gb = (unsigned)((ni + 255) / 256);
    SK(hipMalloc(&flag, 4 * (size_t)nprim));
    SK(hipMalloc(&pos, 4 * (size_t)nprim));
    SK(hipMalloc(&tmp, 4 * scan_tmp_words((uint32_t)nprim)));
    SK(hipMalloc(&total, 4));
    SK(hipMemset(flag, 0, 4 * (size_t)nprim));
    k_sah_mark<<<gb, 256>>>(ni, d_first, d_last, d_split, leafmax, flag);
    SK(hipGetLastError());
    if (int st = scan_exclusive_u32(flag, pos, (uint32_t)nprim, total, tmp, nullptr)) { cleanup(); return st; }
    uint32_t n = 0;
    SK(hipMemcpy(&n, total, 4, hipMemcpyDeviceToHost));
    if (n < 2) { cleanup(); return 1; }
    SK(hipMalloc(&L0, sizeof(SahLeaf) * (size_t)n));
    SK(hipMalloc(&L1, sizeof(SahLeaf) * (size_t)n));
    k_sah_gather<<<gb, 256>>>(ni, d_first, d_last, d_split, leafmax, d_nodes, pos, L0);
    SK(hipGetLastError());
    // the root task: every leaf, its centroid bounds
    SK(hipMalloc(&cb, 24));
    {
        const int init[6] = {0x7f7fffff, 0x7f7fffff, 0x7f7fffff, (int)0x80800000, (int)0x80800000, (int)0x80800000};
        SK(hipMemcpy(cb, init, 24, hipMemcpyHostToDevice));   // lo = max float, hi = ordered -max float
    }
    k_sah_root<<<std::min(1024u, (n + 255) / 256), 256>>>(L0, n, cb);
    const uint32_t cap = n / 2 + 1;
    SK(hipMalloc(&T0, sizeof(SahTask) * (size_t)cap));
    SK(hipMalloc(&T1, sizeof(SahTask) * (size_t)cap));
    k_sah_root_task<<<1, 64>>>(cb, n, T0);
    SK(hipMalloc(&out, sizeof(Node) * (size_t)(n - 1)));
    SK(hipMalloc(&cnt, 4));
    SK(hipMalloc(&dep, 4));
    SK(hipMemset(dep, 0, 4));
    SK(hipGetLastError());
    // tasks of at least kBigTask leaves go across the chip (the top levels); the host reads
    // the task list while it is short to find them.  flag, pos and tmp are free by now and
    // hold the big tasks' flags, their prefix and the scan's scratch.
    constexpr uint32_t kHostTasks = 1024;
    SK(hipMalloc(&big, sizeof(SahBig)));
    SK(hipMalloc(&tsmall, sizeof(SahTask) * kHostTasks));
    std::vector<SahTask> ht, small;
    std::vector<uint32_t> bigs;
    uint32_t ntask = 1;
    SahLeaf *src = L0, *dst = L1;
    SahTask *tin = T0, *tout = T1;
    for (uint32_t level = 0; ntask; level++) {
        if (level > 4096) { cleanup(); sfree(out); return fail(LSG_ERR_INVALID, "sah_rebuild_device: no progress"); }
        SK(hipMemset(cnt, 0, 4));
        bigs.clear();
        small.clear();
        if (ntask <= kHostTasks) {
            ht.resize(ntask);
            SK(hipMemcpy(ht.data(), tin, sizeof(SahTask) * ntask, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < ntask; i++) {
                if (ht[i].end - ht[i].begin >= kBigTask) bigs.push_back(i);
                else small.push_back(ht[i]);
            }
        }
        if (bigs.empty()) {
            k_sah_level<<<ntask, 64>>>(tin, src, dst, out, tout, cnt, dep);
        } else {
            if (!small.empty()) {
                SK(hipMemcpy(tsmall, small.data(), sizeof(SahTask) * small.size(), hipMemcpyHostToDevice));
                k_sah_level<<<(unsigned)small.size(), 64>>>(tsmall, src, dst, out, tout, cnt, dep);
            }
            for (uint32_t i : bigs) {
                const SahTask* t = tin + i;
                const uint32_t m = ht[i].end - ht[i].begin;
                const unsigned g = std::min(2048u, (m + 255) / 256);
                k_big_init<<<1, 256>>>(big);
                k_big_bin<<<g, 256>>>(t, src, big);
                k_big_split<<<1, 64>>>(t, big);
                k_big_flag<<<g, 256>>>(t, src, big, flag);
                if (int st = scan_exclusive_u32(flag, pos, m, nullptr, tmp, nullptr)) { cleanup(); sfree(out); return st; }
                k_big_scatter<<<g, 256>>>(t, src, dst, flag, pos, big);
                k_big_emit<<<1, 64>>>(t, dst, big, out, tout, cnt, dep);
            }
        }
        SK(hipGetLastError());
        SK(hipMemcpy(&ntask, cnt, 4, hipMemcpyDeviceToHost));
        if (ntask > cap) { cleanup(); sfree(out); return fail(LSG_ERR_INVALID, "sah_rebuild_device: %u tasks over %u", ntask, cap); }
        std::swap(src, dst);
        std::swap(tin, tout);
    }
    uint32_t d = 0;
    SK(hipMemcpy(&d, dep, 4, hipMemcpyDeviceToHost));
    cleanup();
#undef SK
    *out_nodes = out; *out_count = n - 1; *out_depth = d;
    return 0;
}

// The BVH4 of a BVH2 (nodes, nn, root >= 0) on the GPU, breadth first: heights bottom-up
// (one launch per level),
// then one level of BVH4 nodes per launch, slots in level order.  budget = the traversal
// stack entries minus one.  Sets C.d_nodes4, C.nnodes4, C.root4 = 0, C.stack_need.
int collapse4_device(ColliderImpl& C, uint32_t budget)
{
    const uint32_t nn = C.nnodes;
    uint32_t *H = nullptr, *cnt = nullptr, *prefix = nullptr, *pnext = nullptr, *tmp = nullptr, *total = nullptr,
             *need = nullptr;
    C4Work *w0 = nullptr, *w1 = nullptr;
    Node4* out = nullptr;
    auto cleanup = [&]() {
        sfree(H); sfree(cnt); sfree(prefix); sfree(pnext); sfree(tmp); sfree(total); sfree(need);
        sfree(w0); sfree(w1);
    };
#define SK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); sfree(out); return hip_fail(e_, #x, __LINE__); } } while (0)
    const unsigned gb = (nn + 255) / 256;
    SK(hipMalloc(&H, 4 * (size_t)nn));
    SK(hipMemset(H, 0, 4 * (size_t)nn));
    uint32_t hroot = 0;
    for (uint32_t k = 1;; k++) {
        k_c4_height<<<gb, 256>>>(nn, C.d_nodes, H, k);
        if (k % 8 == 0) {
            SK(hipMemcpy(&hroot, H + C.root, 4, hipMemcpyDeviceToHost));
            if (hroot) break;
            if (k > 4096) { cleanup(); return fail(LSG_ERR_INVALID, "collapse4: no progress"); }
        }
    }
    SK(hipGetLastError());
    if (hroot > budget) { cleanup(); return fail(LSG_ERR_LIMIT, "collapse4: tree height %u over stack budget %u", hroot, budget); }
    SK(hipMalloc(&out, sizeof(Node4) * (size_t)nn));
    SK(hipMalloc(&w0, sizeof(C4Work) * (size_t)nn));
    SK(hipMalloc(&w1, sizeof(C4Work) * (size_t)nn));
    SK(hipMalloc(&cnt, 4 * (size_t)nn));
    SK(hipMalloc(&prefix, 4 * (size_t)nn));
    SK(hipMalloc(&pnext, 4 * (size_t)nn));
    SK(hipMalloc(&tmp, 4 * scan_tmp_words(nn)));
    SK(hipMalloc(&total, 4));
    SK(hipMalloc(&need, 4));
    SK(hipMemset(need, 0, 4));
    {
        const C4Work r{C.root, 0u, 0u};
        SK(hipMemcpy(w0, &r, sizeof r, hipMemcpyHostToDevice));
    }
    uint32_t nw = 1, n4 = 1;
    while (nw) {
        if (n4 > nn) { cleanup(); sfree(out); return fail(LSG_ERR_INVALID, "collapse4: more BVH4 nodes than BVH2 nodes"); }
        const unsigned g = (nw + 255) / 256;
        k_c4_open<<<g, 256>>>(w0, nw, C.d_nodes, H, budget, out, cnt, pnext, need);
        if (int st = scan_exclusive_u32(cnt, prefix, nw, total, tmp, nullptr)) { cleanup(); sfree(out); return st; }
        k_c4_link<<<g, 256>>>(w0, nw, prefix, pnext, n4, out, w1);
        SK(hipGetLastError());
        uint32_t t = 0;
        SK(hipMemcpy(&t, total, 4, hipMemcpyDeviceToHost));
        n4 += t;
        nw = t;
        std::swap(w0, w1);
    }
    uint32_t sn = 0;
    SK(hipMemcpy(&sn, need, 4, hipMemcpyDeviceToHost));
    Node4* exact = nullptr;
    SK(hipMalloc(&exact, sizeof(Node4) * (size_t)n4));
    SK(hipMemcpy(exact, out, sizeof(Node4) * (size_t)n4, hipMemcpyDeviceToDevice));
    cleanup();
    sfree(out);
#undef SK
    C.d_nodes4 = exact;
    C.bytes += sizeof(Node4) * (uint64_t)n4;
    C.nnodes4 = n4; C.root4 = 0; C.stack_need = sn;
    return LSG_OK;
}

} // namespace lsg
