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
 *
 * The ordered variants (no reference counterpart; k_hgo_*) give the same
 * cache for every schedule: each bucket's new keys inserted one at a time in
 * ascending checksum order, a selection key drawn from the query's contents,
 * and compaction in cell order through per-wave counts.
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
        if (o < kHashGridMaxQueries) {  // the compact lists hold kHashGridMaxQueries entries
            G(out_index)[o] = G(selection)[cell] & ((1u << kHashGridMaxQueriesShift) - 1u);
            G(out_addr)[o] = cell;
        }
        ++o;
    }
}

__global__ void __launch_bounds__(64) k_hg_compact(const uint32_t* selection, uint32_t* out_index, uint32_t* out_addr, uint32_t* out_counter) {
    k_hg_compact_body(selection, out_index, out_addr, out_counter, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Order-independent cache update (vgi_pass_hash_grid_query_select_ordered,
// vgi_pass_hash_grid_query_compact_ordered).  The select runs in four
// launches over the queries / buckets:
//   gather  each query whose key the reference's probe would not find before
//           the bucket's first free slot (a new key, or one whose older copy
//           sits behind a slot the decay pass emptied) adds the key to its
//           bucket's candidate set (the winner buffer, zero on entry, used as
//           a 16-slot table per bucket; the storage is only read)
//   insert  per bucket, the candidates in ascending checksum order through
//           the reference's find-or-insert: the result of the reference's
//           select under the schedule that runs the keys one at a time in
//           that order
//   word    each query's key word, drawn from the query's contents, into
//           the cell's selection word by atomicMax (bit 31 set)
//   claim   the queries whose word equals the cell's maximum write their
//           index into the winner buffer
// ---------------------------------------------------------------------------

// Random selection key of a query: the reference's draw (jenkins of the
// seed, one xorshift step) seeded by a hash of the query's 24 bytes instead
// of its index, with bit 31 set.  Equal queries draw equal words.
__device__ __forceinline__ uint32_t query_select_word(const HashQuery& q, uint32_t frame_hash) {
    uint32_t rng = jenkins_hash(query_content_hash(q) ^ frame_hash);
    return lcg_xorshift(rng) | 0x80000000u;
}

// Threshold word for the next frame's appends: the queries counted under the
// word `thr` were a sample of about n * 2^32 / (2^32 - thr) candidates;
// keep about kHashGridQueryTarget of as many next time (integer arithmetic,
// so every device and the oracle agree).
VGI_HD uint32_t next_query_threshold(uint32_t n, uint32_t thr) {
    const uint64_t keep = (1ull << 32) - thr;
    return query_threshold_for(((uint64_t)n << 32) / keep);
}

// Cell of a key after the insert launch: its first slot in the bucket (the
// cell the reference's find-or-insert returned for it in that schedule), or
// false when the bucket overflowed.
__device__ __forceinline__ bool hash_grid_key_cell(const HashCell* buf, uint32_t hash, uint32_t checksum,
                                                   uint32_t& addr) {
    const uint32_t beg = hash << kHashGridBucketShift;
    for (addr = beg; addr < beg + kHashGridBucketSize; ++addr)
        if (G(&buf[addr].checksum)[0] == checksum) return true;
    return false;
}

__device__ __forceinline__ void k_hgo_gather_body(const HashQuery* queries, const uint32_t* counter,
                                                  const HashCell* storage, uint32_t* cand, uint32_t* threshold,
                                                  uint32_t bx_) {
    const uint32_t id = bx_ * 64u + threadIdx.x;
    if (id == 0 && threshold) G(threshold)[0] = next_query_threshold(G(counter)[0], G(threshold)[0]);
    if (id >= kHashGridMaxQueries || id >= G(counter)[0]) return;
    const HashQuery q = gld(queries + id);
    const uint32_t beg = q.hash << kHashGridBucketShift;
    uint32_t addr = beg;
    for (; addr < beg + kHashGridBucketSize; ++addr) {
        const uint32_t c = G(&storage[addr].checksum)[0];
        if (c == q.checksum) return;  // found in front of every free slot
        if (c == 0u) break;
    }
    if (addr == beg + kHashGridBucketSize) return;  // full bucket without the key: no insert
    for (uint32_t a = beg; a < beg + kHashGridBucketSize; ++a) {
        const uint32_t slot = __hip_atomic_load(G(&cand[a]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (slot == q.checksum) return;
        if (slot == 0u) {
            const uint32_t prev = g_atomic_cas(&cand[a], 0u, q.checksum);
            if (prev == 0u || prev == q.checksum) return;
        }
    }
}

__device__ __forceinline__ void k_hgo_insert_body(HashCell* storage, const uint32_t* cand, uint32_t bx_) {
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    const uint32_t bucket = bx_ * 64u + threadIdx.x;
    if (bucket >= kHashGridNumBuckets) return;
    const uint32_t base = bucket * kHashGridBucketSize;
    const __attribute__((address_space(1))) U4* cw = (const __attribute__((address_space(1))) U4*)(cand + base);
    uint32_t key[kHashGridBucketSize];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const U4 v = cw[k];
        key[4 * k] = v.x;
        key[4 * k + 1] = v.y;
        key[4 * k + 2] = v.z;
        key[4 * k + 3] = v.w;
    }
    if (key[0] == 0u) return;  // candidates fill the table from its first slot
    {
        // leave the table zero for the next frame (the claim launch writes
        // the winner indices into the same buffer, the compaction clears them)
        __attribute__((address_space(1))) U4* cz = (__attribute__((address_space(1))) U4*)(cand + base);
        const U4 z = {0u, 0u, 0u, 0u};
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) cz[k] = z;
    }
    // ascending checksums, the empty slots (0) last
#pragma unroll
    for (uint32_t i = 0; i < kHashGridBucketSize; ++i)
#pragma unroll
        for (uint32_t j = 0; j + 1 < kHashGridBucketSize - i; ++j) {
            const uint32_t a = key[j] - 1u, b = key[j + 1] - 1u;  // 0 wraps to the largest
            const bool sw = a > b;
            const uint32_t lo = sw ? key[j + 1] : key[j], hi = sw ? key[j] : key[j + 1];
            key[j] = lo;
            key[j + 1] = hi;
        }
    uint32_t ck[kHashGridBucketSize];
#pragma unroll
    for (uint32_t k = 0; k < kHashGridBucketSize; ++k) ck[k] = G(&storage[base + k].checksum)[0];
    uint32_t filled = 0;
#pragma unroll
    for (uint32_t i = 0; i < kHashGridBucketSize; ++i) {
        if (key[i] == 0u) break;
        bool done = false;
#pragma unroll
        for (uint32_t s = 0; s < kHashGridBucketSize; ++s) {
            if (done) continue;
            if (ck[s] == key[i]) {
                done = true;
            } else if (ck[s] == 0u) {
                ck[s] = key[i];
                filled |= 1u << s;
                done = true;
            }
        }
    }
#pragma unroll
    for (uint32_t s = 0; s < kHashGridBucketSize; ++s)
        if ((filled >> s) & 1u) G(&storage[base + s].checksum)[0] = ck[s];
}

__device__ __forceinline__ void k_hgo_word_body(const HashQuery* queries, const uint32_t* counter,
                                                const HashCell* storage, uint32_t* selection, uint32_t frame_hash,
                                                uint32_t bx_) {
    const uint32_t id = bx_ * 64u + threadIdx.x;
    if (id >= kHashGridMaxQueries || id >= G(counter)[0]) return;
    const HashQuery q = gld(queries + id);
    uint32_t addr;
    if (!hash_grid_key_cell(storage, q.hash, q.checksum, addr)) return;
    g_atomic_max(&selection[addr], query_select_word(q, frame_hash));
}

__device__ __forceinline__ void k_hgo_claim_body(const HashQuery* queries, const uint32_t* counter,
                                                 const HashCell* storage, const uint32_t* selection, uint32_t* winner,
                                                 uint32_t frame_hash, uint32_t bx_) {
    const uint32_t id = bx_ * 64u + threadIdx.x;
    if (id >= kHashGridMaxQueries || id >= G(counter)[0]) return;
    const HashQuery q = gld(queries + id);
    uint32_t addr;
    if (!hash_grid_key_cell(storage, q.hash, q.checksum, addr)) return;
    if (G(selection)[addr] == query_select_word(q, frame_hash)) G(winner)[addr] = id;
}

// Compaction in cell order: `count` stores each wave's number of selected
// cells (and, wave 0, the counter's value on entry) in the workspace;
// `place` sums the counts of the waves before it and writes its cells there.
// Selected words have bit 31 set.
__device__ __forceinline__ uint64_t compact_lane_mask(const uint32_t* selection, uint32_t first) {
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    const __attribute__((address_space(1))) U4* src = (const __attribute__((address_space(1))) U4*)(selection + first);
    uint64_t act = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCompactCellsPerLane / 4; ++k) {
        const U4 v = src[k];
        act |= (uint64_t)((v.x >> 31) | ((v.y >> 31) << 1) | ((v.z >> 31) << 2) | ((v.w >> 31) << 3)) << (4u * k);
    }
    return act;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) v += __shfl_xor(v, d);
    return v;
}

__device__ __forceinline__ void k_hgo_count_body(const uint32_t* selection, const uint32_t* counter, uint32_t* ws,
                                                 uint32_t bx_) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t act = compact_lane_mask(selection, bx_ * kCompactCellsPerWave + lane * kCompactCellsPerLane);
    const uint32_t total = wave_sum((uint32_t)__popcll(act));
    if (lane == 0) {
        G(ws)[bx_] = total;
        if (bx_ == 0) G(ws)[kCompactWaves] = G(counter)[0];
    }
}

__device__ __forceinline__ void k_hgo_place_body(uint32_t* selection, uint32_t* winner,
                                                 uint32_t* out_index, uint32_t* out_addr, uint32_t* out_counter,
                                                 const uint32_t* ws, uint32_t bx_) {
    static_assert(kCompactWaves == 4 * 64, "four wave counts per lane");
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t w = lane * 4u + k, c = G(ws)[w];
        all += c;
        before += w < bx_ ? c : 0u;
    }
    before = wave_sum(before);
    all = wave_sum(all);
    const uint32_t start = G(ws)[kCompactWaves];
    if (bx_ == kCompactWaves - 1 && lane == 0) G(out_counter)[0] = start + all;
    const uint32_t first = bx_ * kCompactCellsPerWave + lane * kCompactCellsPerLane;
    uint64_t act = compact_lane_mask(selection, first);
    const uint32_t cnt = (uint32_t)__popcll(act);
    uint32_t incl = cnt;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    uint32_t o = start + before + incl - cnt;
    while (act != 0u) {
        const uint32_t k = (uint32_t)__ffsll((long long)act) - 1u;
        act &= act - 1u;
        const uint32_t cell = first + k;
        if (o < kHashGridMaxQueries) {
            G(out_index)[o] = G(winner)[cell] & (kHashGridMaxQueries - 1u);
            G(out_addr)[o] = cell;
        }
        ++o;
        // the selection and winner words go back to zero: the next frame's
        // ordered select needs no clear of the two 4 MiB buffers
        G(selection)[cell] = 0u;
        G(winner)[cell] = 0u;
    }
}

__global__ void __launch_bounds__(64) k_hgo_gather(const HashQuery* queries, const uint32_t* counter,
                                                   const HashCell* storage, uint32_t* cand, uint32_t* threshold) {
    k_hgo_gather_body(queries, counter, storage, cand, threshold, blockIdx.x);
}
__global__ void __launch_bounds__(64) k_hgo_insert(HashCell* storage, const uint32_t* cand) {
    k_hgo_insert_body(storage, cand, blockIdx.x);
}
__global__ void __launch_bounds__(64) k_hgo_word(const HashQuery* queries, const uint32_t* counter,
                                                 const HashCell* storage, uint32_t* selection, uint32_t frame_hash) {
    k_hgo_word_body(queries, counter, storage, selection, frame_hash, blockIdx.x);
}
__global__ void __launch_bounds__(64) k_hgo_claim(const HashQuery* queries, const uint32_t* counter,
                                                  const HashCell* storage, const uint32_t* selection, uint32_t* winner,
                                                  uint32_t frame_hash) {
    k_hgo_claim_body(queries, counter, storage, selection, winner, frame_hash, blockIdx.x);
}
__global__ void __launch_bounds__(64) k_hgo_count(const uint32_t* selection, const uint32_t* counter, uint32_t* ws) {
    k_hgo_count_body(selection, counter, ws, blockIdx.x);
}
__global__ void __launch_bounds__(64) k_hgo_place(uint32_t* selection, uint32_t* winner,
                                                  uint32_t* out_index, uint32_t* out_addr, uint32_t* out_counter,
                                                  const uint32_t* ws) {
    k_hgo_place_body(selection, winner, out_index, out_addr, out_counter, ws, blockIdx.x);
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

// Order-independent select: four launches over one table (the jobs share
// their layout, HgOSelectArgs).
struct HgOSelectArgs {
    const HashQuery* queries;
    const uint32_t* counter;
    HashCell* storage;
    uint32_t *selection, *winner, *threshold;
    uint32_t frame_hash;
};
template <class>
struct JobHgOGather : HgOSelectArgs {
    __device__ static void run(const JobHgOGather& j, uint32_t bx, uint32_t) {
        k_hgo_gather_body(j.queries, j.counter, j.storage, j.winner, j.threshold, bx);
    }
};
template <class>
struct JobHgOInsert : HgOSelectArgs {
    __device__ static void run(const JobHgOInsert& j, uint32_t bx, uint32_t) {
        k_hgo_insert_body(j.storage, j.winner, bx);
    }
};
template <class>
struct JobHgOWord : HgOSelectArgs {
    __device__ static void run(const JobHgOWord& j, uint32_t bx, uint32_t) {
        k_hgo_word_body(j.queries, j.counter, j.storage, j.selection, j.frame_hash, bx);
    }
};
template <class>
struct JobHgOClaim : HgOSelectArgs {
    __device__ static void run(const JobHgOClaim& j, uint32_t bx, uint32_t) {
        k_hgo_claim_body(j.queries, j.counter, j.storage, j.selection, j.winner, j.frame_hash, bx);
    }
};
struct HgOCompactArgs {
    uint32_t *selection, *winner;
    uint32_t *out_index, *out_addr, *out_counter, *ws;
};
template <class>
struct JobHgOCount : HgOCompactArgs {
    __device__ static void run(const JobHgOCount& j, uint32_t bx, uint32_t) {
        k_hgo_count_body(j.selection, j.out_counter, j.ws, bx);
    }
};
template <class>
struct JobHgOPlace : HgOCompactArgs {
    __device__ static void run(const JobHgOPlace& j, uint32_t bx, uint32_t) {
        k_hgo_place_body(j.selection, j.winner, j.out_index, j.out_addr, j.out_counter, j.ws, bx);
    }
};
static_assert(sizeof(JobHgOGather<StorageF32>) == sizeof(HgOSelectArgs) &&
                  sizeof(JobHgOClaim<StorageF32>) == sizeof(HgOSelectArgs) &&
                  sizeof(JobHgOPlace<StorageF32>) == sizeof(HgOCompactArgs),
              "the launches of one table share its layout");

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

static vgi_status fill_hg_oselect(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobHgOGather>(host);
    for (int v = 0; v < n; ++v) {
        const vgi_hash_grid_ordered_select_io& io = recs[v]->io.hg_oselect;
        VGI_REQUIRE(((uintptr_t)io.rw_hash_grid_query_winner_buffer & 15u) == 0,
                    "batched ordered select: winner buffer not 16-byte aligned");
        HgOSelectArgs& a = t->jobs[v];
        a.queries = reinterpret_cast<const HashQuery*>(io.hash_grid_query_buffer);
        a.counter = io.hash_grid_query_counter_buffer;
        a.storage = reinterpret_cast<HashCell*>(io.rw_hash_grid_storage_buffer);
        a.selection = io.rw_hash_grid_query_selection_buffer;
        a.winner = io.rw_hash_grid_query_winner_buffer;
        a.threshold = io.rw_hash_grid_query_threshold;
        a.frame_hash = recs[v]->pc.frame_hash.frame_hash;
        static_assert(kHashGridMaxQueries / 64 >= kHashGridNumBuckets / 64, "the query tiles cover the buckets");
        batch_add_tiles(t->hd, v, kHashGridMaxQueries / 64, 1);
    }
    return VGI_OK;
}
static vgi_status launch_hg_oselect(const void* dev, uint32_t storage, uint32_t grid, size_t lds, hipStream_t st) {
    VGI_CHECK(batch_launch<JobHgOGather>(dev, storage, grid, lds, st));
    VGI_CHECK(batch_launch<JobHgOInsert>(dev, storage, grid, lds, st));
    VGI_CHECK(batch_launch<JobHgOWord>(dev, storage, grid, lds, st));
    return batch_launch<JobHgOClaim>(dev, storage, grid, lds, st);
}
static vgi_status fill_hg_ocompact(const PassRec* const* recs, int n, void* host) {
    auto* t = batch_table<JobHgOCount>(host);
    for (int v = 0; v < n; ++v) {
        const vgi_hash_grid_ordered_compact_io& io = recs[v]->io.hg_ocompact;
        VGI_REQUIRE(((uintptr_t)io.rw_hash_grid_query_selection_buffer & 15u) == 0,
                    "batched ordered compaction: selection buffer not 16-byte aligned");
        HgOCompactArgs& a = t->jobs[v];
        a.selection = io.rw_hash_grid_query_selection_buffer;
        a.winner = io.rw_hash_grid_query_winner_buffer;
        a.out_index = io.rw_compact_query_index_buffer;
        a.out_addr = io.rw_compact_query_cell_addr_buffer;
        a.out_counter = io.rw_compact_query_counter_buffer;
        a.ws = io.rw_workspace_buffer;
        batch_add_tiles(t->hd, v, kCompactWaves, 1);
    }
    return VGI_OK;
}
static vgi_status launch_hg_ocompact(const void* dev, uint32_t storage, uint32_t grid, size_t lds, hipStream_t st) {
    VGI_CHECK(batch_launch<JobHgOCount>(dev, storage, grid, lds, st));
    return batch_launch<JobHgOPlace>(dev, storage, grid, lds, st);
}

const BatchOps* batch_ops_hashgrid(int kind) {
    static const BatchOps decay{sizeof(BatchTable<JobHgDecay<StorageF32>>), fill_hg_decay, batch_launch<JobHgDecay>, 64,
                                false};
    static const BatchOps select{sizeof(BatchTable<JobHgSelect<StorageF32>>), fill_hg_select, batch_launch<JobHgSelect>,
                                 64, false};
    static const BatchOps compact{sizeof(BatchTable<JobHgCompact<StorageF32>>), fill_hg_compact,
                                  batch_launch<JobHgCompact>, 64, false};
    static const BatchOps oselect{sizeof(BatchTable<JobHgOGather<StorageF32>>), fill_hg_oselect, launch_hg_oselect,
                                  64, false};
    static const BatchOps ocompact{sizeof(BatchTable<JobHgOCount<StorageF32>>), fill_hg_ocompact, launch_hg_ocompact,
                                   64, false};
    switch (kind) {
        case kPassHgSelectOrdered: return &oselect;
        case kPassHgCompactOrdered: return &ocompact;
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

extern "C" vgi_status vgi_pass_hash_grid_query_select_ordered(const vgi_hash_grid_ordered_select_io* io,
                                                              const vgi_pc_frame_hash* pc, void* stream) {
    VGI_REQUIRE(io && pc, "vgi_pass_hash_grid_query_select_ordered: null argument");
    VGI_REQUIRE(io->hash_grid_query_buffer && io->hash_grid_query_counter_buffer && io->rw_hash_grid_storage_buffer &&
                    io->rw_hash_grid_query_selection_buffer && io->rw_hash_grid_query_winner_buffer,
                "vgi_pass_hash_grid_query_select_ordered: missing binding");
    VGI_REQUIRE(((uintptr_t)io->rw_hash_grid_query_winner_buffer & 15u) == 0,
                "vgi_pass_hash_grid_query_select_ordered: winner buffer not 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const HashQuery* q = reinterpret_cast<const HashQuery*>(io->hash_grid_query_buffer);
    HashCell* cells = reinterpret_cast<HashCell*>(io->rw_hash_grid_storage_buffer);
    ProfScope prof("hash_grid_query_select_ordered", st, 0.0);
    hipLaunchKernelGGL(k_hgo_gather, dim3(kHashGridMaxQueries / 64), dim3(64), 0, st, q,
                       io->hash_grid_query_counter_buffer, cells, io->rw_hash_grid_query_winner_buffer,
                       io->rw_hash_grid_query_threshold);
    VGI_CHECK(check_launch("k_hgo_gather"));
    hipLaunchKernelGGL(k_hgo_insert, dim3(kHashGridNumBuckets / 64), dim3(64), 0, st, cells,
                       io->rw_hash_grid_query_winner_buffer);
    VGI_CHECK(check_launch("k_hgo_insert"));
    hipLaunchKernelGGL(k_hgo_word, dim3(kHashGridMaxQueries / 64), dim3(64), 0, st, q, io->hash_grid_query_counter_buffer,
                       cells, io->rw_hash_grid_query_selection_buffer, pc->frame_hash);
    VGI_CHECK(check_launch("k_hgo_word"));
    hipLaunchKernelGGL(k_hgo_claim, dim3(kHashGridMaxQueries / 64), dim3(64), 0, st, q,
                       io->hash_grid_query_counter_buffer, cells, io->rw_hash_grid_query_selection_buffer,
                       io->rw_hash_grid_query_winner_buffer, pc->frame_hash);
    return check_launch("k_hgo_claim");
}

extern "C" vgi_status vgi_pass_hash_grid_query_compact_ordered(const vgi_hash_grid_ordered_compact_io* io,
                                                               void* stream) {
    VGI_REQUIRE(io && io->rw_hash_grid_query_selection_buffer && io->rw_hash_grid_query_winner_buffer &&
                    io->rw_compact_query_index_buffer && io->rw_compact_query_cell_addr_buffer &&
                    io->rw_compact_query_counter_buffer && io->rw_workspace_buffer,
                "vgi_pass_hash_grid_query_compact_ordered: missing binding");
    VGI_REQUIRE(((uintptr_t)io->rw_hash_grid_query_selection_buffer & 15u) == 0,
                "vgi_pass_hash_grid_query_compact_ordered: selection buffer not 16-byte aligned");
    static_assert(VGI_HASH_GRID_COMPACT_WORKSPACE_WORDS == kCompactWaves + 1, "workspace size");
    hipStream_t st = (hipStream_t)stream;
    ProfScope prof("hash_grid_query_compact_ordered", st, (double)kHashGridNumCells * 8.0);
    hipLaunchKernelGGL(k_hgo_count, dim3(kCompactWaves), dim3(64), 0, st, io->rw_hash_grid_query_selection_buffer,
                       io->rw_compact_query_counter_buffer, io->rw_workspace_buffer);
    VGI_CHECK(check_launch("k_hgo_count"));
    hipLaunchKernelGGL(k_hgo_place, dim3(kCompactWaves), dim3(64), 0, st, io->rw_hash_grid_query_selection_buffer,
                       io->rw_hash_grid_query_winner_buffer, io->rw_compact_query_index_buffer,
                       io->rw_compact_query_cell_addr_buffer, io->rw_compact_query_counter_buffer,
                       io->rw_workspace_buffer);
    return check_launch("k_hgo_place");
}

static_assert(kHashGridQueryTarget == VGI_HASH_GRID_QUERY_TARGET, "query target");

extern "C" uint32_t vgi_hash_grid_query_threshold_for(uint64_t expected_queries) {
    return query_threshold_for(expected_queries);
}

extern "C" vgi_status vgi_pass_clear_buffer(uint32_t* buffer, uint32_t length, uint32_t value, void* stream) {
    VGI_REQUIRE(buffer || length == 0, "vgi_pass_clear_buffer: null buffer");
    return fill_u32(buffer, length, value, (hipStream_t)stream);
}
