// Resident index object (one per device).
#pragma once

#include "ma_common.hpp"

struct ma_index {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t k = 0;
    int64_t n_keys = 0;
    int64_t n_postings = 0;
    int64_t n_bases = 0;
    int64_t n_contigs = 0;
    int32_t ref_has_n = 0;
    int64_t n_cuckoo = 0;  // keys placed by the cuckoo pass of the hash build
    int32_t keep_dense = 1;  // kmers/offsets resident (0: table + postings only)
    int32_t first_bucket_first = 0;  // lookups read the second bucket only when the first is full
    int32_t compact = 0;             // table holds 8-byte slots (see IndexView)
    int32_t c_m = 0, c_r = 0, c_sb = 0;
    int32_t c_inline = 0;            // compact: one-posting keys hold their position (IndexView)
    uint64_t stash_nb = 0;
    int32_t n_cu = 0;  // compute units of the device (launch sizing)
    uint64_t n_buckets = 0;
    ma::DevBuf kmers;     // u64[n_keys]       (dense-v1)
    ma::DevBuf offsets;   // u64[n_keys + 1]   (dense-v1)
    ma::DevBuf postings;  // u32[2 * n_postings] (dense-v1)
    ma::DevBuf ref2;      // 2-bit reference words
    ma::DevBuf refN;      // N flags, same geometry
    ma::DevBuf coff;      // u64[n_contigs + 1]
    ma::DevBuf table;     // HashSlot[n_buckets * BUCKET_SLOTS], or u64[n_buckets * CBUCKET_SLOTS] compact
    ma::DevBuf stash;     // compact only: HashSlot table of the overflow keys
};

// Per-launch state of the map kernel (its work counter).
struct ma_workspace {
    ma_index *ix = nullptr;
    int64_t max_reads = 0;  // 0 = no limit
    int32_t max_len = 0;
    int32_t persistent = 1;  // 1: persistent grid + work counter; 0: grid sized to the batch
    int32_t chunks_per_wave = 0;  // work chunks per resident wave (sets the claim size); 0 = the tuned knob
    ma::DevBuf counter;
    ma::DevBuf split;  // read-length class sub-batches (MA_TUNE_MAP_CLASS_SPLIT), grow-only
    // forward host loop: per-wave vote tables (grow-only) and their tag clock
    ma::DevBuf fwd_tab, fwd_aux;
    int64_t fwd_entries = 0;
    uint64_t fwd_tag = 0;
};

namespace ma {
IndexView index_view(const ma_index *ix);
// Validate map knobs against the index and the longest read (MA_ERR_ARG + message).
int check_map_params(const ma_index *ix, const ma_params *p, int32_t max_len);
}
