// rd_batch.hip — many CPIs of uneven geometry through the chain with one graph
// launch (rd_batch_* in include/rd_amd.h).
//
// Plan. Each stage groups the CPIs into shape classes — the compile-time
// parameters of its kernel:
//   PFB   taps (4, 8, 16 register-window kernels; any other count generic)
//         and whether the CPI keeps the PFB flags
//   FFT   (log2 length, samples per cycle, all-ones schedule)
//   BF    MFMA K steps = ceil(channels / 16)
//   COV   channels rounded up to 16
//   CFAR  row length (bins)
// A class is one kernel launch over the tiles of all its CPIs: a device job
// table holds each CPI's pointers and sizes, and a tile table maps every tile
// (block) to its job. The beamformer's blocks walk their tiles with a grid
// stride (persistent), so each block derives a CPI's weight digits once; the
// other stages run one block per tile, which keeps their register budget at
// that of the single-CPI kernels. Stages run in chain order; the covariance classes run
// on a forked stream beside CFAR (they only share read-only inputs). The whole
// sequence is captured once into a HIP graph.
//
// Autotuner. Per class, the candidate tile and grid shapes are run on the
// batch's own buffers and timed with HIP events; the fastest is kept. Every
// stage is a pure function of its inputs, so re-running a stage with another
// shape rewrites identical outputs.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "rd_fxp.h"
#include "rd_internal.h"

namespace {

#define RD_TRY(expr)                           \
  do {                                         \
    if ((expr) != hipSuccess) return RD_E_HIP; \
  } while (0)

enum Stage { S_PFB = 0, S_FFT, S_BF, S_COV, S_CFAR_ROWS, S_CFAR_GATHER, S_COUNT };
const char* const kStageName[S_COUNT] = {"pfb", "fft", "bf", "cov", "cfar_rows", "cfar_gather"};

// Events staged per CFAR row in a batch; rows with more are recomputed by the
// gather pass.
constexpr uint32_t kBatchCfarSlot = 64;
// A row emits at most n_bins events plus its summary.
uint32_t batch_cfar_slot(uint32_t n_bins) { return n_bins + 1 < kBatchCfarSlot ? n_bins + 1 : kBatchCfarSlot; }

struct Shape {
  uint32_t a = 0;     // PFB: target lanes per CPI for beat chunking (0: one chunk);
                      // BF: columns per tile; COV: frame splits; CFAR: rows per tile
  uint32_t grid = 0;  // persistent grid size (0: one block per tile)
  uint32_t b = 0;     // CFAR: waves per row
  uint32_t ring = 0;  // BF / COV: the loader-fed kernel (tiles of 128 columns / one bin)
};

struct Class {
  Stage stage;
  uint32_t key = 0;
  std::vector<uint32_t> cpis;
  uint32_t max_ch = 0, max_bt = 0;
  int out_mask = 0;  // BF: outputs shared by every job of the class, -1 if they differ
  Shape shape;
  uint32_t n_tiles = 0;
  void* dev = nullptr;
  size_t dev_cap = 0;
  rd::ClassLaunch L{};
  double ms = -1.0;
  uint32_t tried = 0;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

struct rd_batch {
  rd_ctx* ctx = nullptr;
  uint32_t n = 0;
  std::vector<rd_cpi_desc> d;
  std::vector<rd_cpi_bufs> b;
  std::vector<rd::CfarJob> cfar;  // per CPI
  std::vector<Class> cls;
  uint32_t n_cu = 0;
  void* cfar_scratch = nullptr;
  void* scan_jobs = nullptr;
  uint32_t max_rows = 0;  // CFAR rows of the largest CPI
  uint64_t* totals_dev = nullptr;
  uint64_t* totals_host = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string plan;
};

namespace {

// ------------------------------------------------------------------ tables
// Job and tile tables of a class for a shape. Returns the number of tiles.
template <typename Job>
void put_job(std::vector<uint8_t>& jobs, const Job& j) {
  const size_t o = jobs.size();
  jobs.resize(o + sizeof(Job));
  memcpy(jobs.data() + o, &j, sizeof(Job));
}

uint32_t pfb_chunk(uint32_t P, uint32_t CH, uint32_t F, uint32_t T, uint32_t lanes_target) {
  if (lanes_target == 0) return F;
  const uint64_t lanes = (uint64_t)P * CH;
  const uint64_t want = (lanes_target + lanes - 1) / lanes;
  uint32_t chunk = (uint32_t)((F + want - 1) / want);
  if (chunk < T) chunk = T;
  if (chunk > F) chunk = F;
  return chunk;
}

uint32_t build_tables(const rd_batch& B, const Class& c, const Shape& sh, std::vector<uint8_t>& jobs,
                      std::vector<uint32_t>& tile_job) {
  jobs.clear();
  tile_job.clear();
  uint32_t tile0 = 0;
  for (uint32_t j = 0; j < (uint32_t)c.cpis.size(); ++j) {
    const uint32_t i = c.cpis[j];
    const rd_cpi_desc& d = B.d[i];
    const rd_cpi_bufs& b = B.b[i];
    const uint32_t CH = d.n_chan, N = d.n_fft, F = d.n_frames, NB = d.n_beams;
    uint32_t tiles = 0;
    switch (c.stage) {
      case S_PFB: {
        const uint32_t chunk = pfb_chunk(N, CH, F, d.taps, sh.a);
        const uint32_t nbx = (N + 255) / 256, nz = (F + chunk - 1) / chunk;
        tiles = nbx * CH * nz;
        put_job(jobs, rd::PfbJob{b.iq, b.taps, nullptr, b.u, b.pfb_flags, N, d.taps, CH, F, chunk, nbx, nz, tile0});
        break;
      }
      case S_FFT: {
        const uint32_t log2n = rd::ilog2u(N), used = (1u << log2n) - 1u;
        const uint32_t fpb = rd::fft_frames_per_tile(log2n, (d.scale_sched & used) == used);
        tiles = (CH * F + fpb - 1) / fpb;
        put_job(jobs, rd::FftJob{b.u, nullptr, b.X, b.fft_flags, d.scale_sched, 1u, CH * F, tile0});
        break;
      }
      case S_BF: {
        const uint64_t cols = (uint64_t)F * N;
        tiles = (uint32_t)((cols + sh.a - 1) / sh.a);
        put_job(jobs, rd::BfJob{b.X, b.weights, b.y, b.yflags, b.power, nullptr, b.Xa, cols, CH, NB,
                                (NB + 15) / 16, tile0});
        break;
      }
      case S_COV: {
        const uint32_t nb = rd::cov_bins_per_tile(c.key);
        tiles = (N + nb - 1) / nb;
        put_job(jobs, rd::CovJob{b.Xa, b.R, CH, F, N, tile0});
        break;
      }
      case S_CFAR_ROWS:
      case S_CFAR_GATHER: {
        rd::CfarJob cj = B.cfar[i];
        tiles = (cj.n_rows + sh.a - 1) / sh.a;
        cj.tile0 = tile0;
        put_job(jobs, cj);
        break;
      }
      default: break;
    }
    tile_job.insert(tile_job.end(), tiles, j);
    tile0 += tiles;
  }
  return tile0;
}

rd_status upload(Class& c, const rd_batch& B, const Shape& sh) {
  std::vector<uint8_t> jobs;
  std::vector<uint32_t> tj;
  const uint32_t tiles = build_tables(B, c, sh, jobs, tj);
  const size_t jb = align256(jobs.size()), need = jb + tj.size() * sizeof(uint32_t) + 4;
  if (need > c.dev_cap) {
    if (c.dev) RD_TRY(hipFree(c.dev));
    c.dev = nullptr;
    c.dev_cap = 0;
    RD_TRY(hipMalloc(&c.dev, need));
    c.dev_cap = need;
  }
  RD_TRY(hipMemcpy(c.dev, jobs.data(), jobs.size(), hipMemcpyHostToDevice));
  RD_TRY(hipMemcpy(static_cast<uint8_t*>(c.dev) + jb, tj.data(), tj.size() * sizeof(uint32_t),
                   hipMemcpyHostToDevice));
  c.shape = sh;
  c.n_tiles = tiles;
  c.L.jobs = c.dev;
  c.L.tile_job = reinterpret_cast<const uint32_t*>(static_cast<uint8_t*>(c.dev) + jb);
  c.L.n_tiles = tiles;
  c.L.grid = (sh.grid && sh.grid < tiles) ? sh.grid : tiles;
  if (sh.ring && c.stage == S_BF) c.L.grid = tiles < B.n_cu ? tiles : B.n_cu;  // what the launch uses
  return RD_OK;
}

hipError_t launch_class(const rd_batch& B, const Class& c, hipStream_t s) {
  switch (c.stage) {
    case S_PFB: return rd::launch_pfb_batch(s, c.key & 0xFFu, (c.key >> 8) & 1u, c.L);
    case S_FFT: return rd::launch_fft_batch(s, B.ctx->tw, c.key & 15u, (c.key >> 4) & 15u, (c.key >> 8) & 1u, c.L);
    case S_BF:
      return c.shape.ring ? rd::launch_bf_ring_batch(s, c.key, c.max_bt, B.n_cu, c.out_mask, c.L)
                          : rd::launch_bf_batch(s, c.key, c.shape.a, c.max_ch, c.max_bt, c.out_mask, c.L);
    case S_COV:
      return c.shape.ring ? rd::launch_cov_ring_batch(s, c.key, c.L) : rd::launch_cov_batch(s, c.key, c.shape.a, c.L);
    case S_CFAR_ROWS: return rd::launch_cfar_rows_batch(s, c.key, c.shape.b, c.shape.a, c.L);
    case S_CFAR_GATHER: return rd::launch_cfar_gather_batch(s, c.key, c.shape.b, c.shape.a, c.L);
    default: return hipErrorInvalidValue;
  }
}

// The whole batch in chain order on the context stream, covariance forked.
rd_status enqueue_all(const rd_batch& B) {
  hipStream_t s = B.ctx->stream;
  bool any_cov = false;
  for (const Class& c : B.cls) any_cov |= c.stage == S_COV;
  for (int st = S_PFB; st <= S_BF; ++st)
    for (const Class& c : B.cls)
      if (c.stage == st) RD_TRY(launch_class(B, c, s));
  if (any_cov) {
    RD_TRY(hipEventRecord(B.ev_fork, s));
    RD_TRY(hipStreamWaitEvent(B.side, B.ev_fork, 0));
    for (const Class& c : B.cls)
      if (c.stage == S_COV) RD_TRY(launch_class(B, c, B.side));
    RD_TRY(hipEventRecord(B.ev_join, B.side));
  }
  for (const Class& c : B.cls)
    if (c.stage == S_CFAR_ROWS) RD_TRY(launch_class(B, c, s));
  RD_TRY(rd::launch_cfar_scan_batch(s, B.scan_jobs, B.n, B.max_rows));
  for (const Class& c : B.cls)
    if (c.stage == S_CFAR_GATHER) RD_TRY(launch_class(B, c, s));
  if (any_cov) RD_TRY(hipStreamWaitEvent(s, B.ev_join, 0));
  return RD_OK;
}

// ------------------------------------------------------------------ shapes
bool bf_fits(const Class& c, uint32_t tc) {
  const size_t words = (size_t)c.max_bt * c.key * 64 * 8 + (size_t)c.max_bt * 32 + (size_t)tc * (c.max_ch + 1);
  return words * 4 <= 160 * 1024;
}

// Every CPI of the class fits the loader-fed kernel of its stage.
bool ring_fits(const rd_batch& B, const Class& c) {
  for (uint32_t i : c.cpis) {
    const rd_cpi_desc& d = B.d[i];
    const rd_cpi_bufs& b = B.b[i];
    if (c.stage == S_BF) {
      const uint64_t cols = (uint64_t)d.n_frames * d.n_fft;
      if ((c.key != 3 && c.key != 4) || d.n_chan != 16 * c.key || d.n_beams % 16 != 0 || cols % 128 != 0 ||
          ((uintptr_t)b.X & 15u))
        return false;
    } else if (c.stage == S_COV) {
      if ((c.key != 32 && c.key != 64) || d.n_chan != c.key || d.n_fft % (64 / c.key) != 0 ||
          ((uintptr_t)b.Xa & 15u))
        return false;
    } else {
      return false;
    }
  }
  return !c.cpis.empty();
}

Shape default_shape(const rd_batch& B, const Class& c) {
  Shape s;
  if (ring_fits(B, c)) {
    s.a = c.stage == S_BF ? 128 : 1;
    s.ring = 1;
    return s;
  }
  switch (c.stage) {
    case S_PFB: s.a = 262144; break;
    case S_FFT: break;
    case S_BF:
      s.a = c.key <= 2 ? 256 : 128;
      if (!bf_fits(c, s.a)) s.a = 64;
      s.grid = 8 * B.n_cu;
      break;
    case S_COV: {
      uint32_t min_steps = ~0u;
      for (uint32_t i : c.cpis) min_steps = std::min(min_steps, (B.d[i].n_frames + 31) / 32);  // 32-frame steps
      s.a = std::max(1u, std::min(rd::cov_max_splits(c.key), min_steps));
      break;
    }
    case S_CFAR_ROWS:
    case S_CFAR_GATHER:
      s.b = rd::cfar_min_waves(c.key);
      s.a = rd::cfar_rows_per_tile_max(c.key, s.b);
      break;
    default: break;
  }
  return s;
}

std::vector<Shape> candidates(const rd_batch& B, const Class& c) {
  std::vector<Shape> v;
  const uint32_t cu = B.n_cu;
  switch (c.stage) {
    case S_PFB:
      for (uint32_t a : {0u, 32768u, 65536u, 131072u, 262144u}) v.push_back(Shape{a, 0});
      break;
    case S_FFT:
      v.push_back(Shape{0, 0});
      break;
    case S_BF:
      if (ring_fits(B, c)) v.push_back(Shape{128, 0, 0, 1});
      for (uint32_t tc : {64u, 128u, 256u})
        if (bf_fits(c, tc))
          for (uint32_t g : {0u, 4 * cu, 8 * cu}) v.push_back(Shape{tc, g});
      break;
    case S_COV:
      if (ring_fits(B, c)) v.push_back(Shape{1, 0, 0, 1});
      for (uint32_t f = 1; f <= rd::cov_max_splits(c.key); f *= 2) v.push_back(Shape{f, 0});
      break;
    case S_CFAR_ROWS:
    case S_CFAR_GATHER:
      for (uint32_t w = rd::cfar_min_waves(c.key), k = 0; w <= 8 && k < 3; w *= 2, ++k)
        for (uint32_t r = 1; r <= rd::cfar_rows_per_tile_max(c.key, w); r *= 2) v.push_back(Shape{r, 0, w});
      break;
    default: break;
  }
  return v;
}

// Average time of one launch of class c in its current shape.
rd_status time_class(const rd_batch& B, const Class& c, hipEvent_t e0, hipEvent_t e1, double* ms) {
  hipStream_t s = B.ctx->stream;
  RD_TRY(launch_class(B, c, s));  // warm-up
  const int reps = 3;
  RD_TRY(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) RD_TRY(launch_class(B, c, s));
  RD_TRY(hipEventRecord(e1, s));
  RD_TRY(hipEventSynchronize(e1));
  float t = 0.f;
  RD_TRY(hipEventElapsedTime(&t, e0, e1));
  *ms = (double)t / reps;
  return RD_OK;
}

rd_status autotune(rd_batch& B) {
  // One pass in the default shapes produces every stage's inputs.
  rd_status st = enqueue_all(B);
  if (st != RD_OK) return st;
  RD_TRY(hipStreamSynchronize(B.ctx->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RD_TRY(hipEventCreate(&e0));
  RD_TRY(hipEventCreate(&e1));
  for (Class& c : B.cls) {
    Shape best = c.shape;
    double best_ms = 1e30;
    for (const Shape& sh : candidates(B, c)) {
      st = upload(c, B, sh);
      if (st == RD_OK) {
        double ms = 0;
        st = time_class(B, c, e0, e1, &ms);
        if (st == RD_OK && ms < best_ms) {
          best_ms = ms;
          best = sh;
        }
        ++c.tried;
      }
      if (st != RD_OK) break;
    }
    if (st == RD_OK) st = upload(c, B, best);
    if (st != RD_OK) break;
    c.ms = best_ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

std::string describe(const rd_batch& B) {
  std::string s;
  char line[256];
  snprintf(line, sizeof line, "batch: %u CPIs, %zu class launches + 1 scan\n", B.n, B.cls.size());
  s += line;
  for (const Class& c : B.cls) {
    char key[64];
    if (c.stage == S_FFT)
      snprintf(key, sizeof key, "n=%u spc=%u all1=%u", 1u << (c.key & 15u), (c.key >> 4) & 15u, (c.key >> 8) & 1u);
    else if (c.stage == S_PFB)
      snprintf(key, sizeof key, "taps=%s flags=%u", (c.key & 0xFFu) ? std::to_string(c.key & 0xFFu).c_str() : "any",
               (c.key >> 8) & 1u);
    else if (c.stage == S_BF)
      snprintf(key, sizeof key, "steps=%u", c.key);
    else if (c.stage == S_COV)
      snprintf(key, sizeof key, "chp=%u", c.key);
    else
      snprintf(key, sizeof key, "bins=%u", c.key);
    const char* an = c.stage == S_PFB ? "chunk_lanes" : c.stage == S_BF ? "tile_cols" : c.stage == S_COV ? "frame_splits"
                   : c.stage == S_FFT ? "-" : "rows_per_tile";
    snprintf(line, sizeof line, "%-11s %-22s cpis=%-4zu tiles=%-8u %s=%u grid=%u", kStageName[c.stage], key,
             c.cpis.size(), c.n_tiles, an, c.shape.a, c.L.grid);
    s += line;
    if (c.stage == S_CFAR_ROWS || c.stage == S_CFAR_GATHER) {
      snprintf(line, sizeof line, " waves_per_row=%u", c.shape.b);
      s += line;
    }
    if (c.shape.ring) s += " loader-fed";
    if (c.ms >= 0) {
      snprintf(line, sizeof line, " tuned %.4f ms (%u candidates)", c.ms, c.tried);
      s += line;
    }
    s += "\n";
  }
  return s;
}

void release(rd_batch* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  for (Class& c : B->cls)
    if (c.dev) (void)hipFree(c.dev);
  if (B->cfar_scratch) (void)hipFree(B->cfar_scratch);
  if (B->scan_jobs) (void)hipFree(B->scan_jobs);
  if (B->totals_dev) (void)hipFree(B->totals_dev);
  if (B->totals_host) (void)hipHostFree(B->totals_host);
  if (B->ev_fork) (void)hipEventDestroy(B->ev_fork);
  if (B->ev_join) (void)hipEventDestroy(B->ev_join);
  if (B->side) (void)hipStreamDestroy(B->side);
  delete B;
}

rd_status build(rd_batch& B, uint32_t flags) {
  const uint32_t n = B.n;
  int dev = 0, cus = 0;
  RD_TRY(hipGetDevice(&dev));
  RD_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  B.n_cu = cus > 0 ? (uint32_t)cus : 256;
  RD_TRY(hipStreamCreateWithFlags(&B.side, hipStreamNonBlocking));
  RD_TRY(hipEventCreateWithFlags(&B.ev_fork, hipEventDisableTiming));
  RD_TRY(hipEventCreateWithFlags(&B.ev_join, hipEventDisableTiming));
  RD_TRY(hipMalloc(&B.totals_dev, n * sizeof(uint64_t)));
  RD_TRY(hipHostMalloc(&B.totals_host, n * sizeof(uint64_t), hipHostMallocDefault));

  // CFAR scratch: per CPI [n_rows + 1] offsets then [n_rows][slot][3] staging.
  std::vector<size_t> off(n);
  size_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    off[i] = total;
    const size_t rows = (size_t)B.d[i].n_beams * B.d[i].n_frames;
    total += align256((rows + 1) * 8 + rows * batch_cfar_slot(B.d[i].n_fft) * 24);
  }
  RD_TRY(hipMalloc(&B.cfar_scratch, total));
  B.cfar.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const rd_cpi_desc& d = B.d[i];
    const rd_cpi_bufs& b = B.b[i];
    const uint32_t rows = d.n_beams * d.n_frames;
    B.max_rows = rows > B.max_rows ? rows : B.max_rows;
    uint64_t* offs = reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(B.cfar_scratch) + off[i]);
    B.cfar[i] = rd::CfarJob{b.power, nullptr, offs + rows + 1, b.row_count, offs, b.events, B.totals_dev + i,
                            b.events ? b.events_cap : 0, rd::cfar_geom(d.cfar, d.max_guard, d.max_ref), d.n_fft,
                            rows, d.n_frames, batch_cfar_slot(d.n_fft), 0};
  }
  RD_TRY(hipMalloc(&B.scan_jobs, n * sizeof(rd::CfarJob)));
  RD_TRY(hipMemcpy(B.scan_jobs, B.cfar.data(), n * sizeof(rd::CfarJob), hipMemcpyHostToDevice));

  // Shape classes, in a fixed order per stage (ascending keys).
  std::map<uint32_t, std::vector<uint32_t>> by[S_COUNT];
  for (uint32_t i = 0; i < n; ++i) {
    const rd_cpi_desc& d = B.d[i];
    const uint32_t T = d.taps;
    by[S_PFB][((T == 4 || T == 8 || T == 16) ? T : 0) | (B.b[i].pfb_flags ? 0x100u : 0u)].push_back(i);
    const uint32_t log2n = rd::ilog2u(d.n_fft), used = (1u << log2n) - 1u;
    by[S_FFT][log2n | (d.spc << 4) | (((d.scale_sched & used) == used ? 1u : 0u) << 8)].push_back(i);
    by[S_BF][(d.n_chan + 15) / 16].push_back(i);
    if (B.b[i].R) by[S_COV][(d.n_chan + 15) / 16 * 16].push_back(i);
    by[S_CFAR_ROWS][d.n_fft].push_back(i);
    if (B.b[i].events && B.b[i].events_cap) by[S_CFAR_GATHER][d.n_fft].push_back(i);
  }
  for (int st = 0; st < S_COUNT; ++st)
    for (auto& kv : by[st]) {
      Class c;
      c.stage = (Stage)st;
      c.key = kv.first;
      c.cpis = kv.second;
      bool first = true;
      for (uint32_t i : c.cpis) {
        c.max_ch = std::max(c.max_ch, B.d[i].n_chan);
        c.max_bt = std::max(c.max_bt, (B.d[i].n_beams + 15) / 16);
        const rd_cpi_bufs& bb = B.b[i];
        const int m = (bb.y ? 1 : 0) | (bb.yflags ? 2 : 0) | 4;  // power is always written
        c.out_mask = first ? m : (c.out_mask == m ? m : -1);
        first = false;
      }
      B.cls.push_back(std::move(c));
    }
  for (Class& c : B.cls) {
    rd_status st = upload(c, B, default_shape(B, c));
    if (st != RD_OK) return st;
  }
  if (flags & RD_BATCH_AUTOTUNE) {
    rd_status st = autotune(B);
    if (st != RD_OK) return st;
  }

  hipStream_t s = B.ctx->stream;
  RD_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const rd_status st = enqueue_all(B);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &graph);
  if (st != RD_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return st != RD_OK ? st : RD_E_HIP;
  }
  B.graph = graph;
  RD_TRY(hipGraphInstantiate(&B.exec, graph, nullptr, nullptr, 0));
  B.plan = describe(B);
  return RD_OK;
}

}  // namespace

extern "C" {

rd_status rd_batch_create(rd_ctx* ctx, uint32_t n_cpi, const rd_cpi_desc* descs, const rd_cpi_bufs* bufs,
                          uint32_t flags, rd_batch** out) {
  if (!out) return RD_E_ARG;
  *out = nullptr;
  rd_status st = rd::ctx_enter(ctx);
  if (st != RD_OK) return st;
  if (n_cpi == 0 || n_cpi > 65535 || !descs || !bufs || (flags & ~RD_BATCH_AUTOTUNE)) return RD_E_ARG;
  for (uint32_t i = 0; i < n_cpi; ++i) {
    st = rd::check_cpi(&descs[i]);
    if (st == RD_OK && descs[i].x_layout != 0) st = RD_E_ARG;  // a batch keeps X in natural order
    if (st != RD_OK) return st;
    st = rd::check_cpi_bufs(&bufs[i]);
    if (st != RD_OK) return st;
    if (!bufs[i].Xa) return RD_E_ARG;  // the batch's beamformer classes write the copy its covariance reads
  }
  rd_batch* B = new (std::nothrow) rd_batch();
  if (!B) return RD_E_ARG;
  B->ctx = ctx;
  B->n = n_cpi;
  B->d.assign(descs, descs + n_cpi);
  B->b.assign(bufs, bufs + n_cpi);
  st = build(*B, flags);
  if (st != RD_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    release(B);
    return st;
  }
  ++ctx->live_graphs;
  *out = B;
  return RD_OK;
}

rd_status rd_batch_launch(rd_batch* B) {
  if (!B) return RD_E_ARG;
  rd_status st = rd::ctx_enter(B->ctx);
  if (st != RD_OK) return st;
  return hipGraphLaunch(B->exec, B->ctx->stream) == hipSuccess ? RD_OK : RD_E_HIP;
}

rd_status rd_batch_events(rd_batch* B, uint64_t* n_events_host) {
  if (!B || !n_events_host) return RD_E_ARG;
  rd_status st = rd::ctx_enter(B->ctx);
  if (st != RD_OK) return st;
  RD_TRY(hipMemcpyAsync(B->totals_host, B->totals_dev, B->n * sizeof(uint64_t), hipMemcpyDeviceToHost,
                        B->ctx->stream));
  RD_TRY(hipStreamSynchronize(B->ctx->stream));
  bool over = false;
  for (uint32_t i = 0; i < B->n; ++i) {
    n_events_host[i] = B->totals_host[i];
    over |= B->b[i].events && B->totals_host[i] > B->b[i].events_cap;
  }
  return over ? RD_E_CAPACITY : RD_OK;
}

rd_status rd_batch_plan(const rd_batch* B, char* text, size_t cap) {
  if (!B || !text || cap == 0) return RD_E_ARG;
  const size_t k = std::min(cap - 1, B->plan.size());
  memcpy(text, B->plan.data(), k);
  text[k] = 0;
  return B->plan.size() < cap ? RD_OK : RD_E_CAPACITY;
}

rd_status rd_batch_destroy(rd_batch* B) {
  if (!B) return RD_E_ARG;
  (void)hipSetDevice(B->ctx->device);
  (void)hipStreamSynchronize(B->ctx->stream);
  (void)hipStreamSynchronize(B->side);
  --B->ctx->live_graphs;
  release(B);
  return RD_OK;
}

}  // extern "C"
