// Segmented work queue for persistent launches.
//
// The entries [0, total) of a launch are cut into chunks of kClaimChunk
// consecutive entries, and the chunks are dealt round-robin to up to
// kClaimSegs segments (chunk j to segment j mod nseg), each with its own
// cursor on its own 256-byte line of the work buffer.  A device-scope atomic
// on one line serialises at about 11 ns on MI355X (88 M claims/s), a rate
// that the per-lane refills of a persistent launch reach; cursors on
// separate lines are served in parallel (64 lines: about 4 G/s).  Because
// the segments interleave, the whole chip still works through the queue as
// one band, in queue order, as with a single cursor: the rays in flight are
// neighbours in the queue (the same scene, the same image tiles), which is
// what keeps their nodes in the L2 caches.
//
// A wave takes each claim from the next segment in turn (skipping exhausted
// ones), so the segments empty at the same pace and the band stays narrow;
// a wave that stayed on one segment would let the segments of slow waves fall
// behind.  Which wave traces which entry changes nothing in the results:
// every entry is claimed exactly once and written at its own index.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace wf {

constexpr uint32_t kClaimSegs = 64;                       // most segments (bits of a 64-bit mask)
constexpr uint32_t kClaimStride = 64;                     // words between cursors (256 bytes)
constexpr uint32_t kClaimChunk = 64;                      // consecutive entries dealt to one segment
constexpr uint32_t kWorkClaimBase = 64;                   // first cursor's word in the work buffer
constexpr uint32_t kWorkWords = kWorkClaimBase + kClaimSegs * kClaimStride;

struct SegQueue {
    uint32_t* cur;      // cursor of segment k at cur[k * kClaimStride], zero at launch
    uint64_t total;     // entries in the queue
    uint32_t nseg;      // segments in use: a power of two, at most kClaimSegs
    uint32_t seg;       // the segment the wave claims from next
    uint32_t rseg;      // the segment of the wave's last claim
    uint32_t seen;      // the cursor after the wave's last claim (any segment)
    uint64_t empty;     // segments the wave knows to be exhausted (bit k)

    // wave: the wave's index in the grid; max_segs (a power of two) caps
    // the segment count, which also leaves every segment two chunks or more.
    __device__ void init(uint32_t* cursors, uint64_t n, uint32_t wave, uint32_t max_segs)
    {
        cur = cursors;
        total = n;
        uint32_t ns = 1u;
        while (ns < max_segs && (uint64_t)ns * 4u * kClaimChunk <= n) ns *= 2u;
        nseg = ns;
        empty = n == 0 ? ~0ull : (ns == 64u ? 0ull : ~0ull << ns);
        seen = 0u;
        seg = wave & (ns - 1u);
        rseg = seg;
    }

    // entries of segment k: its whole chunks, and the partial last chunk of
    // the queue when that chunk is dealt to k
    __device__ uint32_t len_of(uint32_t k) const
    {
        if (k >= nseg) return 0u;
        const uint64_t chunks = total / kClaimChunk;
        const uint32_t rest = (uint32_t)(total % kClaimChunk);
        const uint32_t extra = (uint32_t)(chunks & (nseg - 1u));   // chunks % nseg
        const uint64_t whole = chunks / nseg + (k < extra ? 1u : 0u);
        return (uint32_t)(whole * kClaimChunk) + (k == extra ? rest : 0u);
    }

    // queue index of position p of the segment of the last claim
    __device__ uint64_t entry(uint32_t p) const
    {
        return ((uint64_t)(p / kClaimChunk) * nseg + rseg) * kClaimChunk + p % kClaimChunk;
    }

    // entries of the next segment not yet claimed as far as this wave knows
    // (the segments advance together, so any segment's cursor estimates it)
    __device__ uint32_t left_hint() const
    {
        const uint32_t len = len_of(seg);
        return len > seen ? len - seen : 0u;
    }

    // Reads every cursor (one vector load, lane k reads segment k) and moves
    // to the first segment at or after the current one that still has
    // entries.  False when none has.  Wave-uniform; every lane must be active.
    __device__ bool rescan(uint32_t lane)
    {
        uint32_t v = 0xFFFFFFFFu;
        if (lane < nseg && !((empty >> lane) & 1ull))
            v = __hip_atomic_load(cur + lane * kClaimStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t open = __ballot(v < len_of(lane));
        empty = ~open;
        if (open == 0ull) return false;
        const uint64_t rot = seg ? (open >> seg) | (open << (64u - seg)) : open;
        seg = (seg + (uint32_t)__builtin_ctzll(rot)) & 63u;
        return true;
    }

    // Claims up to g positions of the next segment with entries left, with
    // one atomic from the leader lane; returns how many (0: the queue is
    // drained) and the first in `first`; entry() maps them to queue indices.
    // Wave-uniform; every lane must be active.
    __device__ uint32_t claim(uint32_t g, uint32_t lane, uint32_t leader, uint32_t& first)
    {
        for (;;) {
            if ((empty >> seg) & 1ull) {
                if (!rescan(lane)) return 0u;
            }
            uint32_t b = 0u;
            if (lane == leader) b = atomicAdd(cur + seg * kClaimStride, g);
            b = __builtin_amdgcn_readlane(b, leader);   // wave-uniform (scalar registers)
            const uint32_t len = len_of(seg);
            if (b < len) {
                const uint32_t got = len - b < g ? len - b : g;
                if (b + got >= len) empty |= 1ull << seg;
                seen = b + got;
                first = b;
                rseg = seg;
                seg = (seg + 1u) & (nseg - 1u);
                return got;
            }
            empty |= 1ull << seg;
        }
    }
};

} // namespace wf
