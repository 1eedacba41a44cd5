// Marching cubes with wavefront prefix-sum compaction.
//
// The reference launches one 12-thread block per cube and appends triangles
// through a global atomic counter into a buffer of max(R³/2, 1e6) triangles
// (TurboMC.cu:383-478, 572-638), so only 12 of 64 lanes work and the output
// order is unordered.  Here one lane owns one cube, a block a brick of
// 8 × 8 × 4 cubes, and the launch runs twice: a count pass writes one
// triangle count per brick, a scan turns brick counts into offsets, and an
// emit pass skips the bricks without triangles, recomputes the cube case of
// the others, takes its in-brick offset from a block prefix sum (wavefront
// scans) and writes triangles in brick order, cube order inside a brick.
// The result is exact-size and deterministic.  Corner convention, edge
// interpolation, table and the negated face normal follow TurboMC.cu:399-470.
#include "mc.h"
#include "mc_table.h"
#include "scan.h"

namespace dwg {

__constant__ uint64_t c_mc_rows[256];
static bool g_mc_table_ready[64] = {false};

static void ensure_table(dwg_ctx* ctx) {
    if (ctx->device >= 0 && ctx->device < 64 && g_mc_table_ready[ctx->device]) return;
    static const dwg_mc::PackedTable tab = dwg_mc::make_packed_table();
    DWG_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_mc_rows), tab.row, sizeof(tab.row)));
    if (ctx->device >= 0 && ctx->device < 64) g_mc_table_ready[ctx->device] = true;
}

constexpr int kMcThreads = 256;

struct McArgs {
    const float* f;
    int rx, ry, rz;
    int64_t ncubes;
    float ox, oy, oz, h, iso;
};

// the 4 corners of face z of the cubes in column (x, y): face corner j is
// at (corner_dx(j), corner_dy(j)); a cube's corners 0-3 lie on its face
// z + 1, corners 4-7 on its face z
__device__ __forceinline__ void load_face(const McArgs& A, int x, int y, int z, float (&f)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
        f[j] = A.f[int64_t(x + dwg_mc::corner_dx(j)) +
                   int64_t(A.rx) * (int64_t(y + dwg_mc::corner_dy(j)) + int64_t(A.ry) * z)];
}

__device__ __forceinline__ int cube_case(const McArgs& A, const float (&lo)[4], const float (&hi)[4], float (&v)[8]) {
    int c = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = hi[j];
        v[4 + j] = lo[j];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) c |= (v[i] > A.iso) ? (1 << i) : 0;
    return c;
}

// A block takes a brick of 8 × 8 × 16 cubes, each thread a column of 4
// cubes along z whose shared faces it loads once.  Bricks are numbered x,
// y, z fastest first and dealt to the XCDs in contiguous runs
// (xcd_block), so the rows two neighbouring bricks share come from one L2.
// (Bricks of whole 64-cube rows read the grid in longer runs but leave a
// wave few lanes with triangles to write: the emit pass doubles.)
constexpr int kBx = 8, kBy = 8, kZPer = 4, kBz = 4 * kZPer;
static_assert(kBx * kBy * (kBz / kZPer) == kMcThreads, "one column of kZPer cubes per thread");

struct McBricks {
    int nbx, nby, nbz;
};

__device__ __forceinline__ int64_t brick_of_block() { return xcd_block(blockIdx.x, gridDim.x); }

__device__ __forceinline__ void column_of(const McBricks& B, int64_t b, int& x, int& y, int& z0) {
    x = int(b % B.nbx) * kBx + (threadIdx.x & (kBx - 1));
    y = int((b / B.nbx) % B.nby) * kBy + ((threadIdx.x / kBx) & (kBy - 1));
    z0 = int(b / (int64_t(B.nbx) * B.nby)) * kBz + (threadIdx.x / (kBx * kBy)) * kZPer;
}

// the cases of a thread's column; 0 past the grid
__device__ __forceinline__ void column_cases(const McArgs& A, int x, int y, int z0, int (&cases)[kZPer]) {
    float lo[4] = {0.f, 0.f, 0.f, 0.f};
    const bool col = x < A.rx - 1 && y < A.ry - 1;
    if (col && z0 < A.rz - 1) load_face(A, x, y, z0, lo);
#pragma unroll
    for (int j = 0; j < kZPer; ++j) {
        cases[j] = 0;
        if (col && z0 + j < A.rz - 1) {
            float hi[4], v[8];
            load_face(A, x, y, z0 + j + 1, hi);
            cases[j] = cube_case(A, lo, hi, v);
#pragma unroll
            for (int k = 0; k < 4; ++k) lo[k] = hi[k];
        }
    }
}

__device__ __forceinline__ int wave_inclusive_sum(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
    }
    return v;
}

// exclusive prefix of v over the block and the block total
__device__ int block_exclusive_sum(int v, int& total) {
    __shared__ int wsum[kMcThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int incl = wave_inclusive_sum(v);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kMcThreads / 64; ++w) {
        before += w < wave ? wsum[w] : 0;
        total += wsum[w];
    }
    return before + incl - v;
}

// the nt triangles of the cube (x, y, z) with corner values v and table row
// `row`, written from slot o on
__device__ __forceinline__ void emit_cube(const McArgs& A, int x, int y, int z, const float (&v)[8], uint64_t row,
                                          int nt, int64_t o0, float* __restrict__ tris, float* __restrict__ cent,
                                          float* __restrict__ nrm) {
    for (int t = 0; t < nt; ++t) {
        float P[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int e = int((row >> (4 * (3 * t + k))) & 0xF);
            const int a = dwg_mc::edge_a(e), b = dwg_mc::edge_b(e);
            const float pa[3] = {__builtin_fmaf(float(x + dwg_mc::corner_dx(a)), A.h, A.ox),
                                 __builtin_fmaf(float(y + dwg_mc::corner_dy(a)), A.h, A.oy),
                                 __builtin_fmaf(float(z + dwg_mc::corner_dz(a)), A.h, A.oz)};
            const float pb[3] = {__builtin_fmaf(float(x + dwg_mc::corner_dx(b)), A.h, A.ox),
                                 __builtin_fmaf(float(y + dwg_mc::corner_dy(b)), A.h, A.oy),
                                 __builtin_fmaf(float(z + dwg_mc::corner_dz(b)), A.h, A.oz)};
            // P = P1 − ell (P2 − P1), ell = (a1 − iso)/(a2 − a1 + 1e-10)
            const float ell = (v[a] - A.iso) / (v[b] - v[a] + 1e-10f);
#pragma unroll
            for (int d = 0; d < 3; ++d) P[k][d] = pa[d] - ell * (pb[d] - pa[d]);
        }
        const int64_t o = o0 + t;
        if (tris) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int d = 0; d < 3; ++d) tris[9 * o + 3 * k + d] = P[k][d];
        }
        if (cent) {
#pragma unroll
            for (int d = 0; d < 3; ++d) cent[3 * o + d] = (P[0][d] + P[1][d] + P[2][d]) / 3.0f;
        }
        if (nrm) {
            const float u[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
            const float w[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
            const float c[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
            const float s = -1.0f / (sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) + 1e-8f);
#pragma unroll
            for (int d = 0; d < 3; ++d) nrm[3 * o + d] = c[d] * s;
        }
    }
}

// The count pass: one wave per brick, each lane a whole column of kBz cubes
// (its kBz + 1 z faces loaded once), the brick's total a wave sum
constexpr int kCountThreads = kBx * kBy;
static_assert(kCountThreads == 64, "one wave per brick");

__global__ __launch_bounds__(kCountThreads) void mc_count_kernel(McArgs A, McBricks B, int64_t* __restrict__ tile_counts,
                                                                  int* __restrict__ list,
                                                                  unsigned long long* __restrict__ nlist) {
    __shared__ uint8_t s_ntri[256];  // triangles of each case, read per lane from LDS
#pragma unroll
    for (int c = threadIdx.x; c < 256; c += kCountThreads) s_ntri[c] = uint8_t(c_mc_rows[c] >> 60);
    __syncthreads();
    const int64_t b = brick_of_block();
    const int x = int(b % B.nbx) * kBx + (threadIdx.x & (kBx - 1));
    const int y = int((b / B.nbx) % B.nby) * kBy + int(threadIdx.x) / kBx;
    const int z0 = int(b / (int64_t(B.nbx) * B.nby)) * kBz;
    int cnt = 0;
    if (x < A.rx - 1 && y < A.ry - 1 && z0 < A.rz - 1) {
        float lo[4], hi[4], v[8];
        load_face(A, x, y, z0, lo);
#pragma unroll
        for (int j = 0; j < kBz; ++j) {
            if (z0 + j < A.rz - 1) {
                load_face(A, x, y, z0 + j + 1, hi);
                cnt += s_ntri[cube_case(A, lo, hi, v)];
#pragma unroll
                for (int k = 0; k < 4; ++k) lo[k] = hi[k];
            }
        }
    }
    const int incl = wave_inclusive_sum(cnt);
    if (threadIdx.x == 63) {
        tile_counts[b] = incl;
        // the bricks with triangles, in any order: the emit pass takes one
        // block each (its output place comes from the scan of the counts)
        if (incl > 0) list[atomicAdd(nlist, 1ull)] = int(b);
    }
}

__global__ __launch_bounds__(kMcThreads) void mc_emit_kernel(McArgs A, McBricks B, const int64_t* __restrict__ tile_offsets,
                                                             const int* __restrict__ list, float* __restrict__ tris,
                                                             float* __restrict__ cent, float* __restrict__ nrm) {
    const int64_t b = list[blockIdx.x];  // a brick with triangles
    int x, y, z0;
    column_of(B, b, x, y, z0);
    int cases[kZPer];
    column_cases(A, x, y, z0, cases);
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < kZPer; ++j) cnt += int(c_mc_rows[cases[j]] >> 60);
    int total;
    int64_t o = tile_offsets[b] + block_exclusive_sum(cnt, total);
    if (cnt == 0) return;
    for (int j = 0; j < kZPer; ++j) {  // the column's cubes in z order, each in table order
        const uint64_t row = c_mc_rows[cases[j]];
        const int nt = int(row >> 60);
        if (nt == 0) continue;
        float lo[4], hi[4], v[8];  // its corners again: few cubes have triangles
        load_face(A, x, y, z0 + j, lo);
        load_face(A, x, y, z0 + j + 1, hi);
        (void)cube_case(A, lo, hi, v);
        emit_cube(A, x, y, z0 + j, v, row, nt, o, tris, cent, nrm);
        o += nt;
    }
}

// One pass for callers that take the triangles in any order (the
// diffusion's centroids and face normals, whose sums do not depend on it):
// the count pass's wave per brick also writes the brick's triangles, at a
// range taken with one atomic add once the brick's count is known, each
// lane reloading the corners of the cubes of its column that have
// triangles.  Triangles come brick by brick in the order the waves take
// their ranges, each brick's by column and then in z and table order; their
// values are those of the ordered passes.  A brick writes only when its
// whole range fits in capacity; the count is complete either way.
__global__ __launch_bounds__(kCountThreads) void mc_fused_kernel(McArgs A, McBricks B,
                                                                  unsigned long long* __restrict__ ntri,
                                                                  int64_t capacity, float* __restrict__ cent,
                                                                  float* __restrict__ nrm) {
    __shared__ uint8_t s_ntri[256];
#pragma unroll
    for (int c = threadIdx.x; c < 256; c += kCountThreads) s_ntri[c] = uint8_t(c_mc_rows[c] >> 60);
    __syncthreads();
    const int64_t b = brick_of_block();
    const int x = int(b % B.nbx) * kBx + (threadIdx.x & (kBx - 1));
    const int y = int((b / B.nbx) % B.nby) * kBy + int(threadIdx.x) / kBx;
    const int z0 = int(b / (int64_t(B.nbx) * B.nby)) * kBz;
    int cnt = 0;
    uint32_t has = 0;  // bit j: cube z0 + j of the column has triangles
    if (x < A.rx - 1 && y < A.ry - 1 && z0 < A.rz - 1) {
        float lo[4], hi[4], v[8];
        load_face(A, x, y, z0, lo);
#pragma unroll
        for (int j = 0; j < kBz; ++j) {
            if (z0 + j < A.rz - 1) {
                load_face(A, x, y, z0 + j + 1, hi);
                const int n = s_ntri[cube_case(A, lo, hi, v)];
                cnt += n;
                has |= (n > 0 ? 1u : 0u) << j;
#pragma unroll
                for (int k = 0; k < 4; ++k) lo[k] = hi[k];
            }
        }
    }
    const int incl = wave_inclusive_sum(cnt);
    const int total = __shfl(incl, 63, 64);
    if (total == 0) return;
    uint32_t base_lo = 0, base_hi = 0;
    if (threadIdx.x == 63) {
        const unsigned long long at = atomicAdd(ntri, (unsigned long long)total);
        base_lo = uint32_t(at);
        base_hi = uint32_t(at >> 32);
    }
    const int64_t base = int64_t((uint64_t(uint32_t(__shfl(int(base_hi), 63, 64))) << 32) |
                                 uint32_t(__shfl(int(base_lo), 63, 64)));
    if ((!cent && !nrm) || base + total > capacity) return;
    int64_t o = base + (incl - cnt);
    while (has) {
        const int j = __builtin_ctz(has);
        has &= has - 1u;
        float lo[4], hi[4], v[8];
        load_face(A, x, y, z0 + j, lo);
        load_face(A, x, y, z0 + j + 1, hi);
        const uint64_t row = c_mc_rows[cube_case(A, lo, hi, v)];
        const int nt = int(row >> 60);
        emit_cube(A, x, y, z0 + j, v, row, nt, o, nullptr, cent, nrm);
        o += nt;
    }
}

static McArgs mc_args(const float* field, const dwg_grid* g, float iso, McBricks& B, int64_t& tiles) {
    DWG_REQUIRE(g && g->res[0] >= 2 && g->res[1] >= 2 && g->res[2] >= 2, "grid needs at least 2 samples per axis");
    McArgs A{field, g->res[0], g->res[1], g->res[2], 0, g->origin[0], g->origin[1], g->origin[2], g->h, iso};
    A.ncubes = int64_t(A.rx - 1) * (A.ry - 1) * (A.rz - 1);
    B = McBricks{(A.rx - 1 + kBx - 1) / kBx, (A.ry - 1 + kBy - 1) / kBy, (A.rz - 1 + kBz - 1) / kBz};
    tiles = int64_t(B.nbx) * B.nby * B.nbz;
    DWG_REQUIRE(tiles < (int64_t(1) << 31), "grid too large for the marching-cubes launch");
    return A;
}

int64_t marching_cubes_any_order(dwg_ctx* ctx, const float* field, const dwg_grid* g, float iso, int64_t capacity,
                                 float* cent, float* nrm) {
    DWG_REQUIRE(capacity >= 0, "negative capacity");
    McBricks B;
    int64_t tiles;
    const McArgs A = mc_args(field, g, iso, B, tiles);
    ensure_table(ctx);
    auto* n = static_cast<unsigned long long*>(ctx->buf("mc_fused_count", 64));
    fill_u32(ctx, n, 0u, 2);
    {
        ProfileScope ps(ctx, "mc_fused");
        hipLaunchKernelGGL(mc_fused_kernel, dim3(unsigned(tiles)), dim3(kCountThreads), 0, ctx->stream, A, B, n,
                           capacity, cent, nrm);
        check_launch();
    }
    const int64_t total = int64_t(read_scalar<unsigned long long>(ctx, n));
    if (total > capacity) throw Error(DWG_ERR_CAPACITY, "marching cubes output exceeds capacity", total);
    return total;
}

int64_t marching_cubes(dwg_ctx* ctx, const float* field, const dwg_grid* g, float iso, int64_t capacity, float* tris,
                       float* cent, float* nrm) {
    McBricks B;
    int64_t tiles;
    const McArgs A = mc_args(field, g, iso, B, tiles);
    ensure_table(ctx);
    // brick counts (scanned in place: offsets, the total at [tiles]) and, at
    // [tiles + 1], the number of bricks with triangles
    int64_t* offs = static_cast<int64_t*>(ctx->buf("mc_tiles", sizeof(int64_t) * (tiles + 2)));
    int* list = static_cast<int*>(ctx->buf("mc_list", sizeof(int) * tiles));
    auto* nlist = reinterpret_cast<unsigned long long*>(offs + tiles + 1);
    fill_u32(ctx, nlist, 0u, 2);
    {
        ProfileScope ps(ctx, "mc_count");
        hipLaunchKernelGGL(mc_count_kernel, dim3(unsigned(tiles)), dim3(kCountThreads), 0, ctx->stream, A, B, offs,
                           list, nlist);
        check_launch();
        exclusive_scan<int64_t>(ctx, offs, tiles);
    }
    struct Totals {
        int64_t tris, bricks;
    };
    const Totals tot = read_scalar<Totals>(ctx, reinterpret_cast<const Totals*>(offs + tiles));
    const int64_t total = tot.tris;
    if (total > capacity) throw Error(DWG_ERR_CAPACITY, "marching cubes output exceeds capacity", total);
    if (total > 0 && (tris || cent || nrm)) {
        ProfileScope ps(ctx, "mc_emit");
        hipLaunchKernelGGL(mc_emit_kernel, dim3(unsigned(tot.bricks)), dim3(kMcThreads), 0, ctx->stream, A, B, offs,
                           list, tris, cent, nrm);
        check_launch();
    }
    return total;
}

}  // namespace dwg
