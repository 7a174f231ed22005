// rd_internal.h — launch interface between the C ABI and the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rd_amd.h"

namespace rd {

// Streaming stages (rd_stream.hip)
hipError_t launch_pfb(hipStream_t s, const uint32_t* x, const uint32_t* coeff, uint32_t P, uint32_t T,
                      uint32_t n_chan, uint32_t n_beats, uint32_t* hist, uint32_t* y, uint8_t* flags);
hipError_t launch_fft(hipStream_t s, const uint32_t* tw4096, const uint32_t* in, uint32_t n, uint32_t spc,
                      uint32_t sched, int reorder, uint32_t n_frames, uint32_t* out, uint32_t* bins,
                      uint32_t* stage_flags);
hipError_t launch_align(hipStream_t s, const uint32_t* X, uint32_t n_chan, uint64_t cols, uint32_t* Xa);

// MFMA contractions (rd_mfma.hip)
size_t bf_prep_bytes(uint32_t n_ant, uint32_t n_beams);
hipError_t launch_bf(hipStream_t s, const uint32_t* Xa, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                     uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, int64_t* acc, void* prep);
// The beamformer reading the FFT's channel-major X[ch][col] and writing the
// aligned copy Xa[col][ch] (nullable) from the same staged tiles.
hipError_t launch_bf_chmajor(hipStream_t s, const uint32_t* X, const uint32_t* W, uint32_t n_ant, uint32_t n_beams,
                             uint64_t n_cols, uint32_t* y, uint8_t* flags, uint32_t* power, uint32_t* xa_out);
hipError_t launch_cov(hipStream_t s, const uint32_t* Xa, uint32_t n_chan, uint32_t n_frames, uint32_t n_bins,
                      int64_t* R);

// CFAR (rd_cfar.hip). scratch must hold cfar_scratch_bytes(n_rows).
size_t cfar_scratch_bytes(uint32_t n_rows);
hipError_t launch_cfar(hipStream_t s, const void* cells, uint32_t cell_bytes, uint32_t n_bins, uint32_t n_rows,
                       const rd_cfar_cfg& cfg, uint32_t max_guard, uint32_t max_ref, const uint32_t* frame_ids,
                       uint32_t frame_mod, uint64_t* events, uint64_t events_cap, uint32_t* row_count,
                       void* scratch, uint64_t* total_dev);

}  // namespace rd
