// Paired-end flags and mate fields of one pair of merged map records, shared by
// the PE sort keys (sort.hip) and the PE BAM pack (bam.hip).
//
// Records lie R1 block first ([0, np)) then R2 block ([np, 2 np)); pair p's
// records are p and np + p, emitted as output records 2p (R1) and 2p + 1 (R2).
// Follows the PE branch of _emit_bam_from_batch, linear_gpu_fm.mojo:826-870:
// flag | 1 | 64 (R1) or | 1 | 128 (R2); an index contig outside the target table
// gives tid -1 and flag | 4; a mate without a target sets 0x8 on the other; both
// on one target: 0x2 on both, mate tid / pos, tlen = |pos2 - pos1| + len(R2),
// positive on the leftmost mate (pos1 <= pos2 -> R1); else each mapped mate's
// tid / pos goes to the other's mate fields.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ma {

struct PeMate {
    int64_t flag, tid, pos, ntid, npos, tlen;
};

__device__ __forceinline__ void pe_pair(int64_t np, int64_t p, const int32_t *__restrict__ cid,
                                        const uint32_t *__restrict__ pos, const int32_t *__restrict__ rflag,
                                        const uint32_t *__restrict__ lens, const int32_t *__restrict__ tid_of_cid,
                                        int64_t n_tid, PeMate m[2]) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int64_t r = k ? np + p : p;
        const int64_t c = cid[r];
        m[k].flag = (int64_t)rflag[r] | 1 | (k ? 128 : 64);
        m[k].tid = -1;
        if (c >= 0 && c < n_tid)
            m[k].tid = tid_of_cid[c];
        else
            m[k].flag |= 4;
        m[k].pos = (int64_t)pos[r];
        m[k].ntid = -1;
        m[k].npos = -1;
        m[k].tlen = 0;
    }
    if (m[0].tid < 0) m[1].flag |= 8;
    if (m[1].tid < 0) m[0].flag |= 8;
    if (m[0].tid >= 0 && m[0].tid == m[1].tid) {
        m[0].flag |= 2;
        m[1].flag |= 2;
        m[0].ntid = m[1].ntid = m[0].tid;
        m[0].npos = m[1].pos;
        m[1].npos = m[0].pos;
        const int64_t d = m[1].pos - m[0].pos;
        const int64_t span = (d < 0 ? -d : d) + (int64_t)lens[np + p];
        m[0].tlen = m[0].pos <= m[1].pos ? span : -span;
        m[1].tlen = -m[0].tlen;
    } else {
        if (m[0].tid >= 0) {
            m[1].ntid = m[0].tid;
            m[1].npos = m[0].pos;
        }
        if (m[1].tid >= 0) {
            m[0].ntid = m[1].tid;
            m[0].npos = m[1].pos;
        }
    }
}

}  // namespace ma
