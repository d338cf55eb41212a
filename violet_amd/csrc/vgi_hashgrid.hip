/*
 * vgi_hashgrid.hip — world-space radiance cache maintenance passes:
 *
 *   k_hg_decay    restir/hash_grid_decay.hlsl:13-128   (one lane per bucket)
 *   k_hg_select   restir/hash_grid_query_select.hlsl:22-56 (one lane per query)
 *   k_hg_compact  restir/hash_grid_query_compact.hlsl:78-118
 *
 * Storage: 65536 buckets x 16 cells of {u32 checksum, f32 rgb} (16 MiB),
 * decay 1 byte per cell (read four cells per 32-bit word).  Insertion uses a
 * CAS on the checksum and the per-cell query choice an atomicMax, as in the
 * reference; which query index / which slot wins depends on scheduling, so
 * only the resulting sets are deterministic.  Compaction is a wave-level
 * stream compaction: one global atomic per wave of 4096 cells.
 */
#include "vgi_batch_kernel.h"
#include "vgi_internal.h"
#include "vgi_radiance.h"

namespace vgi {

__device__ __forceinline__ void k_hg_decay_body(uint32_t* decay_words, HashCell* storage, uint32_t bx_, uint32_t by_) {
    const uint32_t bucket = bx_ * 64u + threadIdx.x;
    if (bucket >= kHashGridNumBuckets) return;
    const uint32_t addr_offset = bucket * 16u;
    uint32_t occ = 0;
    uint32_t slice[4] = {0u, 0u, 0u, 0u};
    uint32_t set = 0;
    for (; set < 4u;) {
        uint32_t set_addr = addr_offset + set * 4u;
        uint32_t packed = G(decay_words)[set_addr >> 2];
        if (packed == 0u) break;
        uint32_t updated = 0u;
        uint32_t decay = 0u;
        for (uint32_t i = 0; i < 4u; ++i) {
            uint32_t shift = i * 8u;
            decay = (packed >> shift) & 0xFFu;
            if (decay == 0u) break;
            decay = decay - 1u;
            if (decay == 0u) {
                HashCell e;
                e.checksum = 0u;
                e.r = e.g = e.b = 0.0f;
                gst(storage + set_addr + i, e);
            } else {
                updated |= (decay << shift);
                occ |= (1u << (set * 4u + i));
            }
        }
        slice[set] = updated;
        ++set;
        if (decay == 0u) break;  // (also taken when the set's last cell just expired)
    }
    if (occ != 0u) {
        while (true) {
            uint32_t last = 31u - (uint32_t)__clz(occ);
            uint32_t first_space = (uint32_t)__ffs(~occ) - 1u;
            if (first_space < last) {
                uint32_t src_set = last >> 2, dst_set = first_space >> 2;
                uint32_t last_shift = (last & 3u) * 8u;
                uint32_t d = slice[src_set] >> last_shift;
                slice[src_set] ^= (d << last_shift);
                uint32_t fs_shift = (first_space & 3u) * 8u;
                slice[dst_set] |= (d << fs_shift);
                HashCell c = gld(storage + addr_offset + last);
                gst(storage + addr_offset + first_space, c);
                HashCell e;
                e.checksum = 0u;
                e.r = e.g = e.b = 0.0f;
                gst(storage + addr_offset + last, e);
                occ |= (1u << first_space);
                occ ^= (1u << last);
            } else {
                break;
            }
        }
    }
    for (uint32_t i = 0; i < set; ++i) G(decay_words)[(addr_offset >> 2) + i] = slice[i];
}

__global__ void __launch_bounds__(64) k_hg_decay(uint32_t* decay_words, HashCell* storage) {
    k_hg_decay_body(decay_words, storage, blockIdx.x, blockIdx.y);
}

// hash_grid.inc.hlsl:149-184
__device__ __forceinline__ bool hash_grid_find_or_insert(HashCell* buf, uint32_t hash, uint32_t checksum,
                                                         uint32_t& addr) {
    uint32_t beg = hash << kHashGridBucketShift;
    for (addr = beg; addr < beg + kHashGridBucketSize; ++addr) {
        uint32_t slot = __hip_atomic_load(G(&buf[addr].checksum), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (slot == checksum) return true;
        if (slot == 0u) {
            uint32_t prev = g_atomic_cas(&buf[addr].checksum, 0u, checksum);
            if (prev == 0u || prev == checksum) return true;
        }
    }
    return false;
}

__device__ __forceinline__ void k_hg_select_body(const HashQuery* queries, const uint32_t* counter, HashCell* storage, uint32_t* selection, uint32_t frame_hash, uint32_t bx_, uint32_t by_) {
    uint32_t id = bx_ * 64u + threadIdx.x;
    if (id >= kHashGridMaxQueries || id >= *counter) return;
    HashQuery q = gld(queries + id);
    uint32_t addr;
    if (!hash_grid_find_or_insert(storage, q.hash, q.checksum, addr)) return;
    uint32_t rng = jenkins_hash(id ^ frame_hash);
    uint32_t r = lcg_xorshift(rng);
    uint32_t word = (r & (0xFFFFFFFFu << (kHashGridMaxQueriesShift + 1u))) | (1u << 31) | id;
    g_atomic_max(&selection[addr], word);
}

__global__ void __launch_bounds__(64) k_hg_select(const HashQuery* queries, const uint32_t* counter, HashCell* storage, uint32_t* selection, uint32_t frame_hash) {
    k_hg_select_body(queries, counter, storage, selection, frame_hash, blockIdx.x, blockIdx.y);
}

// One wave compacts 4096 cells: each lane scans 64 consecutive selection
// words (16-byte loads) into a bit mask, a wave prefix sum over the lanes'
// counts places them, and one atomic per wave reserves the output range
// (a few percent of the cells are selected, so one atomic per 64 cells
// serialises on the counter instead).  Output is in cell order within a wave.
constexpr uint32_t kCompactCellsPerLane = 64;
constexpr uint32_t kCompactCellsPerWave = 64 * kCompactCellsPerLane;
constexpr uint32_t kCompactWaves = kHashGridNumCells / kCompactCellsPerWave;
static_assert(kHashGridNumCells % kCompactCellsPerWave == 0, "compaction covers the cells exactly");

__device__ __forceinline__ void k_hg_compact_body(const uint32_t* selection, uint32_t* out_index, uint32_t* out_addr, uint32_t* out_counter, uint32_t bx_, uint32_t by_) {
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t first = bx_ * kCompactCellsPerWave + lane * kCompactCellsPerLane;
    const __attribute__((address_space(1))) U4* src = (const __attribute__((address_space(1))) U4*)(selection + first);
    uint64_t act = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCompactCellsPerLane / 4; ++k) {
        const U4 v = src[k];
        act |= (uint64_t)((v.x != 0u ? 1u : 0u) | (v.y != 0u ? 2u : 0u) | (v.z != 0u ? 4u : 0u) | (v.w != 0u ? 8u : 0u))
               << (4u * k);
    }
    const uint32_t cnt = (uint32_t)__popcll(act);
    uint32_t incl = cnt;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const uint32_t total = __shfl(incl, 63);
    if (total == 0u) return;
    uint32_t base = 0;
    if (lane == 0) base = g_atomic_add(out_counter, total);
    base = __shfl(base, 0);
    uint32_t o = base + incl - cnt;
    while (act != 0u) {
        const uint32_t k = (uint32_t)__ffsll((long long)act) - 1u;
        act &= act - 1u;
        const uint32_t cell = first + k;
        G(out_index)[o] = G(selection)[cell] & ((1u << kHashGridMaxQueriesShift) - 1u);
        G(out_addr)[o] = cell;
        ++o;
    }
}

__global__ void __launch_bounds__(64) k_hg_compact(const uint32_t* selection, uint32_t* out_index, uint32_t* out_addr, uint32_t* out_counter) {
    k_hg_compact_body(selection, out_index, out_addr, out_counter, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Fused multi-view launches (vgi_batch.h); every view has its own cache.
// ---------------------------------------------------------------------------
template <class>
struct JobHgDecay {
    uint32_t* decay_words;
    HashCell* storage;
    __device__ static void run(const JobHgDecay& j, uint32_t bx, uint32_t by) {
        k_hg_decay_body(j.decay_words, j.storage, bx, by);
    }
};
template <class>
struct JobHgSelect {
    const HashQuery* queries;
    const uint32_t* counter;
    HashCell* storage;
    uint32_t* selection;
    uint32_t frame_hash;
    __device__ static void run(const JobHgSelect& j, uint32_t bx, uint32_t by) {
        k_hg_select_body(j.queries, j.counter, j.storage, j.selection, j.frame_hash, bx, by);
    }
};
template <class>
struct JobHgCompact {
    const uint32_t* selection;
    uint32_t *out_index, *out_addr, *out_counter;
    __device__ static void run(const JobHgCompact& j, uint32_t bx, uint32_t by) {
        k_hg_compact_body(j.selection, j.out_index, j.out_addr, j.out_counter, bx, by);
    }
};

static vgi_status fill_hg_decay(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobHgDecay>(host);
    for (int v = 0; v < n; ++v) {
        const vgi_hash_grid_decay_io& io = recs[v]->io.hg_decay;
        VGI_REQUIRE(((uintptr_t)io.rw_decay_buffer & 3u) == 0, "batched hash-grid decay: unaligned decay buffer");
        t->jobs[v] = {(uint32_t*)io.rw_decay_buffer, reinterpret_cast<HashCell*>(io.rw_storage_buffer)};
        batch_add_tiles(t->hd, v, kHashGridNumBuckets / 64, 1);
    }
    return VGI_OK;
}
static vgi_status fill_hg_select(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobHgSelect>(host);
    for (int v = 0; v < n; ++v) {
        const vgi_hash_grid_query_select_io& io = recs[v]->io.hg_select;
        t->jobs[v] = {reinterpret_cast<const HashQuery*>(io.hash_grid_query_buffer), io.hash_grid_query_counter_buffer,
                      reinterpret_cast<HashCell*>(io.rw_hash_grid_storage_buffer), io.rw_hash_grid_query_selection_buffer,
                      recs[v]->pc.frame_hash.frame_hash};
        batch_add_tiles(t->hd, v, kHashGridMaxQueries / 64, 1);
    }
    return VGI_OK;
}
static vgi_status fill_hg_compact(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobHgCompact>(host);
    for (int v = 0; v < n; ++v) {
        const vgi_hash_grid_query_compact_io& io = recs[v]->io.hg_compact;
        t->jobs[v] = {io.hash_grid_query_selection_buffer, io.rw_compact_query_index_buffer,
                      io.rw_compact_query_cell_addr_buffer, io.rw_compact_query_counter_buffer};
        VGI_REQUIRE(((uintptr_t)io.hash_grid_query_selection_buffer & 15u) == 0,
                    "batched hash-grid compaction: selection buffer not 16-byte aligned");
        batch_add_tiles(t->hd, v, kCompactWaves, 1);
    }
    return VGI_OK;
}

const BatchOps* batch_ops_hashgrid(int kind) {
    static const BatchOps decay{sizeof(BatchTable<JobHgDecay<StorageF32>>), fill_hg_decay, batch_launch<JobHgDecay>, 64,
                                false};
    static const BatchOps select{sizeof(BatchTable<JobHgSelect<StorageF32>>), fill_hg_select, batch_launch<JobHgSelect>,
                                 64, false};
    static const BatchOps compact{sizeof(BatchTable<JobHgCompact<StorageF32>>), fill_hg_compact,
                                  batch_launch<JobHgCompact>, 64, false};
    switch (kind) {
        case kPassHgDecay: return &decay;
        case kPassHgSelect: return &select;
        case kPassHgCompact: return &compact;
        default: return nullptr;
    }
}

}  // namespace vgi

using namespace vgi;

extern "C" vgi_status vgi_pass_hash_grid_decay(const vgi_hash_grid_decay_io* io, void* stream) {
    VGI_REQUIRE(io && io->rw_decay_buffer && io->rw_storage_buffer, "vgi_pass_hash_grid_decay: missing binding");
    VGI_REQUIRE(((uintptr_t)io->rw_decay_buffer & 3u) == 0, "vgi_pass_hash_grid_decay: decay buffer not 4-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("hash_grid_decay", st, (double)kHashGridNumCells * 2.0);
    hipLaunchKernelGGL(k_hg_decay, dim3(kHashGridNumBuckets / 64), dim3(64), 0, st, (uint32_t*)io->rw_decay_buffer,
                       reinterpret_cast<HashCell*>(io->rw_storage_buffer));
    return check_launch("k_hg_decay");
}

extern "C" vgi_status vgi_pass_hash_grid_query_select(const vgi_hash_grid_query_select_io* io, const vgi_pc_frame_hash* pc,
                                                      void* stream) {
    VGI_REQUIRE(io && pc, "vgi_pass_hash_grid_query_select: null argument");
    VGI_REQUIRE(io->hash_grid_query_buffer && io->hash_grid_query_counter_buffer && io->rw_hash_grid_storage_buffer &&
                    io->rw_hash_grid_query_selection_buffer,
                "vgi_pass_hash_grid_query_select: missing binding");
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("hash_grid_query_select", st, 0.0);
    hipLaunchKernelGGL(k_hg_select, dim3(kHashGridMaxQueries / 64), dim3(64), 0, st,
                       reinterpret_cast<const HashQuery*>(io->hash_grid_query_buffer), io->hash_grid_query_counter_buffer,
                       reinterpret_cast<HashCell*>(io->rw_hash_grid_storage_buffer),
                       io->rw_hash_grid_query_selection_buffer, pc->frame_hash);
    return check_launch("k_hg_select");
}

extern "C" vgi_status vgi_pass_hash_grid_query_compact(const vgi_hash_grid_query_compact_io* io, void* stream) {
    VGI_REQUIRE(io && io->hash_grid_query_selection_buffer && io->rw_compact_query_index_buffer &&
                    io->rw_compact_query_cell_addr_buffer && io->rw_compact_query_counter_buffer,
                "vgi_pass_hash_grid_query_compact: missing binding");
    VGI_REQUIRE(((uintptr_t)io->hash_grid_query_selection_buffer & 15u) == 0,
                "vgi_pass_hash_grid_query_compact: selection buffer not 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("hash_grid_query_compact", st, (double)kHashGridNumCells * 4.0);
    hipLaunchKernelGGL(k_hg_compact, dim3(kCompactWaves), dim3(64), 0, st, io->hash_grid_query_selection_buffer,
                       io->rw_compact_query_index_buffer, io->rw_compact_query_cell_addr_buffer,
                       io->rw_compact_query_counter_buffer);
    return check_launch("k_hg_compact");
}

extern "C" vgi_status vgi_pass_clear_buffer(uint32_t* buffer, uint32_t length, uint32_t value, void* stream) {
    VGI_REQUIRE(buffer || length == 0, "vgi_pass_clear_buffer: null buffer");
    return fill_u32(buffer, length, value, (hipStream_t)stream);
}
