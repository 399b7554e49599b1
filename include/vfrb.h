/*
 * vfrb.h - C ABI of the MI355X (gfx950) VLITE FRB search chain.
 *
 * One shared library, libvfrb.so, exports every function declared here.
 * Plain C types only: device buffers are `void*`/typed pointers obtained from
 * vfrb_malloc (or any HIP allocation), streams are opaque `void*` (a
 * hipStream_t; NULL = the legacy default stream).  Every call returns a
 * vfrb_status (0 = success) and never exits the process; the message of the
 * last failure on the calling thread is returned by vfrb_last_error().
 *
 * Each entry point names the reference code it replaces (paths relative to
 * the paulray/vlite-frb tree):
 *   PIPELINE   = vdifFILE2dedisp_PIPELINE.cu
 *   TIMESERIES = vdifDADA2timeseries.cu
 *   STATS      = vdifDADA2stats.cu
 *
 * Threading: plans are not thread safe; use one plan per host thread.  All
 * device work is enqueued asynchronously on the stream argument.
 */
#ifndef VFRB_H
#define VFRB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
typedef int32_t vfrb_status;
#define VFRB_OK              0
#define VFRB_ERR_INVALID     1  /* bad argument or inconsistent shape        */
#define VFRB_ERR_UNSUPPORTED 2  /* shape the kernels do not implement        */
#define VFRB_ERR_NOMEM       3  /* device or host allocation failed          */
#define VFRB_ERR_HIP         4  /* HIP runtime / launch error                */
#define VFRB_ERR_NODEVICE    5  /* no usable GPU                             */
#define VFRB_ERR_IO          6  /* file I/O error                            */
#define VFRB_ERR_FORMAT      7  /* VDIF stream violates the expected layout  */
#define VFRB_ERR_EOF         8  /* end of input reached                      */
#define VFRB_ERR_CAPACITY    9  /* request exceeds a fixed capacity          */

/* Replaces HandleError()'s message text (PIPELINE:124-132). */
const char* vfrb_strerror(vfrb_status status);
/* Message of the last failure on this thread ("" if none). */
const char* vfrb_last_error(void);
const char* vfrb_version(void);

/* ------------------------------------------------------------------------ */
/* Derived parameters (host)                                                */
/* ------------------------------------------------------------------------ */
/* VLITE constants, PIPELINE:33-47. */
#define VFRB_FRAMELEN     5032
#define VFRB_HDRLEN       32
#define VFRB_DATALEN      5000
#define VFRB_FRAMESPERSEC 25600
#define VFRB_SAMPLERATE   128e6
#define VFRB_BW_MHZ       64.0
#define VFRB_LOWFREQ_MHZ  320.0
#define VFRB_MUOSLOW_MHZ  360.0

typedef struct vfrb_params {
    int32_t twonchan;  /* FFT length (2*NCHAN)                             */
    int32_t nchan;     /* NCHAN                                             */
    int32_t nfreq;     /* NCHAN+1 (R2C bins)                                */
    int32_t nave;      /* spectra averaged per output sample                */
    int32_t nbundle;   /* output samples per host bundle                    */
    int32_t nloop;     /* bytes per pol per bundle = NBUNDLE*NAVE*TWONCHAN  */
    int32_t nchanout;  /* channels kept                                     */
    int32_t nskip;     /* first FFT bin kept (channel 0 = highest freq)     */
    int32_t maxnt;     /* rows per dedispersion window                      */
    int32_t nhalfcyc;  /* 128-row bandpass blocks per window                */
    double chanbw;     /* MHz                                               */
    double foff;       /* -chanbw (SIGPROC)                                 */
    double fch1;       /* centre of channel 0, MHz                          */
    double tchan;      /* output sample time, s                             */
    double tbundle;    /* bundle duration, s                                */
    double maxtime;    /* window length, s                                  */
} vfrb_params;

/*
 * Parameter derivation of PIPELINE:289-309 (same formulae in
 * TIMESERIES:276-305, STATS:209-232).  fullband=0 keeps the reference's
 * MUOS cut (MUOSLOWFREQ = 360 MHz); fullband=1 keeps all NCHAN channels
 * (NCHANOUT = NCHAN, NSKIP = 1), the shape of the 1024- and 4096-channel
 * workloads.  nchanout_override > 0 forces NCHANOUT (TIMESERIES:282 uses
 * 3840) with NSKIP = NCHAN - NCHANOUT + 1.
 */
vfrb_status vfrb_params_derive(int32_t twonchan, int32_t nave, int32_t nbundle,
                               double maxtime, int32_t fullband,
                               int32_t nchanout_override, vfrb_params* out);

/*
 * Dispersion delay table, FP64 then truncated to int (bit-exact with the
 * reference formula, TIMESERIES:449-464; the same text is commented at
 * PIPELINE:446-461):
 *   off[i][0] = 0,
 *   off[i][j] = (int)(4.148808e3*DM[i]*(pow(fch1-j*chanbw,-2)-pow(fch1,-2))/tchan)
 *   max_delay = same formula at j = nchanout for the LAST DM of the list.
 * offsets may be NULL (only max_delay wanted); max_delay may be NULL.
 * min_margin (optional) receives the smallest distance of dt/tchan to an
 * integer over the table, i.e. how far any delay is from flipping.
 */
vfrb_status vfrb_delay_table(const float* dms, int32_t ndm, int32_t nchanout,
                             double fch1, double chanbw, double tchan,
                             int32_t* offsets, int32_t* max_delay,
                             double* min_margin);

/* ------------------------------------------------------------------------ */
/* Device runtime plumbing (thin wrappers over HIP)                         */
/* ------------------------------------------------------------------------ */
vfrb_status vfrb_device_count(int32_t* count);
vfrb_status vfrb_set_device(int32_t device);
vfrb_status vfrb_device_name(char* buf, int32_t cap);
vfrb_status vfrb_mem_info(size_t* free_bytes, size_t* total_bytes);
vfrb_status vfrb_malloc(void** ptr, size_t bytes);
vfrb_status vfrb_free(void* ptr);
vfrb_status vfrb_host_alloc(void** ptr, size_t bytes);  /* pinned */
vfrb_status vfrb_host_free(void* ptr);
vfrb_status vfrb_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
vfrb_status vfrb_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
vfrb_status vfrb_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
vfrb_status vfrb_memset(void* dst, int32_t value, size_t bytes, void* stream);
vfrb_status vfrb_stream_create(void** stream);
vfrb_status vfrb_stream_destroy(void* stream);
vfrb_status vfrb_stream_synchronize(void* stream);
vfrb_status vfrb_device_synchronize(void);
vfrb_status vfrb_event_create(void** event);
vfrb_status vfrb_event_destroy(void* event);
vfrb_status vfrb_event_record(void* event, void* stream);
vfrb_status vfrb_event_synchronize(void* event);
vfrb_status vfrb_event_elapsed_ms(float* ms, void* start, void* stop);
/* Work queued on `stream` after this call waits for `event` (cross-stream order). */
vfrb_status vfrb_stream_wait_event(void* stream, void* event);
/*
 * hipGraph capture of any sequence of library calls (no counterpart in the
 * reference, whose loop launches every kernel one by one):
 * vfrb_graph_begin starts capturing the work later enqueued on `stream` (a
 * stream of vfrb_stream_create, not the null stream); vfrb_graph_end stops
 * it and instantiates the graph; vfrb_graph_launch replays it on a stream;
 * vfrb_graph_destroy frees it.  A replay repeats the captured kernels with
 * the captured arguments, so it suits calls whose pointers and sizes do not
 * change.  The library's kernels clear their own counters with kernels, so
 * a captured chain holds no memset node; vfrb_memset and synchronous copies
 * do not belong in a capture.
 */
typedef struct vfrb_graph vfrb_graph;
vfrb_status vfrb_graph_begin(void* stream);
vfrb_status vfrb_graph_end(void* stream, vfrb_graph** graph);
vfrb_status vfrb_graph_launch(vfrb_graph* graph, void* stream);
vfrb_status vfrb_graph_destroy(vfrb_graph* graph);

/* ------------------------------------------------------------------------ */
/* Front end: fused unpack + R2C FFT + detection + pol sum + time average   */
/* ------------------------------------------------------------------------ */
typedef struct vfrb_fb_plan vfrb_fb_plan;

/*
 * Replaces cufftPlan1d(&plan, TWONCHAN, CUFFT_R2C, NAVE*NBUNDLE)
 * (PIPELINE:409).  Power-of-two twonchan in [256, 16384] runs the
 * specialised radix-8 kernel; other even lengths whose factors are in
 * {2,3,4,5,8} (e.g. the reference default 12500) run the mixed-radix kernel.
 */
vfrb_status vfrb_fb_plan_create(vfrb_fb_plan** plan, int32_t twonchan,
                                int32_t nave, int32_t nskip, int32_t nchanout);
vfrb_status vfrb_fb_plan_destroy(vfrb_fb_plan* plan);

/*
 * Replaces, per bundle, convertarray x2 + cufftExecR2C x2 + sumpower +
 * avepower + the NBUNDLE channel-select copies (PIPELINE:582-632,
 * 1292-1322).  pol0/pol1 hold nrows*nave*twonchan 8-bit offset-binary
 * samples each (device).  Output row r goes to
 *   fb[((row0 + r) mod ring_rows) * nchanout + c],  c < nchanout,
 * with fb[.][c] = (1/nave) * sum_frames (|X0[nskip+c]|^2+|X1[nskip+c]|^2)/nfreq.
 * ring_rows = 0 means a linear buffer (no wrap).  Power-of-two lengths read
 * the pols as 32-bit words: pol0 and pol1 must be 4-byte aligned.
 */
vfrb_status vfrb_fb_execute(vfrb_fb_plan* plan, const uint8_t* pol0,
                            const uint8_t* pol1, int64_t nrows, float* fb,
                            int64_t ring_rows, int64_t row0, void* stream);
/*
 * The same filterbank read straight from raw frame pairs through an ingest
 * slot table (power-of-two FFT lengths): row r's bytes are payload-stream
 * bytes byte0 + r*nave*twonchan onwards (byte0 a multiple of 4), zero slots
 * reading as zero bytes.  No de-framed copy is made.
 */
vfrb_status vfrb_fb_execute_frames(vfrb_fb_plan* plan, const uint8_t* frames,
                                   const int32_t* table, int64_t nslots, int64_t byte0,
                                   int64_t nrows, float* fb, int64_t ring_rows,
                                   int64_t row0, void* stream);

/* ------------------------------------------------------------------------ */
/* Bandpass (time-major, in place)                                          */
/* ------------------------------------------------------------------------ */
/*
 * subtractbpCAL (PIPELINE:1427-1474; flat variant TIMESERIES:1365-1407):
 * for every channel and every block of 128 rows (last block may be short):
 * mean, "sig" = variance (sic), clipped mean over (mean-3sig, mean+3sig),
 * subtract the clipped mean.  fb is f32 [nrows][nch], device, in place.
 */
vfrb_status vfrb_bandpass_inplace(float* fb, int32_t nrows, int32_t nch, void* stream);

/*
 * Per-channel statistics of STATS calcstats (STATS:910-946): per channel and
 * block of ncolstatblk rows, mean, variance (named sig), and the outlier
 * fraction computed with integer division (count/npts).  Outputs are f32
 * [nstatblks][nch].
 */
vfrb_status vfrb_chan_stats(const float* fb, int32_t nrows, int32_t nch,
                            int32_t ncolstatblk, int32_t nstatblks,
                            float* ave, float* sig, float* outlier, void* stream);

/*
 * subtractbpCAL followed by the per-channel statistics, as the STATS program
 * runs them on every window (STATS:412-416), in one read of the raw
 * filterbank: replaces vfrb_bandpass_inplace + vfrb_chan_stats with the same
 * results bit for bit, and leaves fb unchanged.
 */
vfrb_status vfrb_bandpass_chan_stats(const float* fb, int32_t nrows, int32_t nch,
                                     int32_t ncolstatblk, int32_t nstatblks,
                                     float* ave, float* sig, float* outlier, void* stream);

/* ------------------------------------------------------------------------ */
/* Device-side frame ingest (no reference counterpart: the reference copies   */
/* payloads frame by frame on the host, PIPELINE:934-1171)                    */
/* ------------------------------------------------------------------------ */
/*
 * Raw VDIF frame pairs already in device memory (a NIC-to-GPU path): pair k
 * is frames 2k (thread 0 = pol 0) and 2k+1 (thread 1 = pol 1), 5032 bytes
 * each, the first pair being the one the reference's initframe01 +
 * initskip01 start on (thread-aligned, cal boundary).  vfrb_ingest_frames
 * validates every pair and builds the slot table of the two payload streams
 * with the reference's zero-filled runs for missing frames (checkframeskip,
 * PIPELINE:888-930; seconds-aware rule TIMESERIES:893-937 when
 * seconds_check): table[s] = pair feeding payload slot s (5000 bytes per pol)
 * or -1 for a zero slot.  The parse state (frame number expected next)
 * carries from one call to the next, so a stream may arrive in batches, each
 * with its own table starting at slot 0.  A pair whose threads are not 0/1
 * or whose frame numbers (seconds) differ is VFRB_ERR_FORMAT: the host
 * sequencer reproduces the reference on such streams.  A call that fails
 * (VFRB_ERR_FORMAT, or VFRB_ERR_CAPACITY when the slots exceed table_cap)
 * leaves the parse state as it was: a retry with a larger table continues
 * from the same expected frame.  Synchronizes `stream` (the slot count is
 * returned to the host).
 */
typedef struct vfrb_ingest vfrb_ingest;
vfrb_status vfrb_ingest_create(vfrb_ingest** ingest, int64_t max_pairs, int32_t seconds_check);
vfrb_status vfrb_ingest_destroy(vfrb_ingest* ingest);
/* Expected next frame number and second; -1 = accept the next pair as is. */
vfrb_status vfrb_ingest_reset(vfrb_ingest* ingest, int32_t expected_frame, int32_t expected_second);
vfrb_status vfrb_ingest_frames(vfrb_ingest* ingest, const uint8_t* frames, int64_t npairs,
                               int32_t* table, int64_t table_cap, int64_t* nslots, void* stream);
/*
 * No counterpart in the reference.  vfrb_ingest_frames for callers that
 * cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it allocates nothing.  Its
 * scratch is a file-scope device array of 12.6 MB per GPU, held until the
 * process ends; at most 2^20 frame pairs per call (larger counts are
 * VFRB_ERR_CAPACITY).  The parse state lives with the caller:
 * *expected_frame / *expected_second are the frame number and second
 * expected next (-1 = accept the first pair as is) and receive the state
 * after the batch; a failed call leaves them unchanged.  Same table and
 * errors as vfrb_ingest_frames; synchronizes `stream`.  Calls that overlap
 * on one GPU (several host threads or streams) share the scratch and race.
 */
vfrb_status vfrb_ingest_frames_noalloc(const uint8_t* frames, int64_t npairs, int32_t seconds_check,
                                       int32_t* expected_frame, int32_t* expected_second,
                                       int32_t* table, int64_t table_cap, int64_t* nslots,
                                       void* stream);
/* The de-framed payload of one pol: nslots*5000 bytes (zero slots as zeros). */
vfrb_status vfrb_ingest_gather(const uint8_t* frames, const int32_t* table, int64_t nslots,
                               int32_t pol, uint8_t* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Back end: search plan (replaces the dedisp plan)                         */
/* ------------------------------------------------------------------------ */
#define VFRB_STORE_F32 0  /* reference: f32 residuals                        */
#define VFRB_STORE_F16 1  /* f16 residuals, window-global power-of-2 scale   */
#define VFRB_STORE_U8  2  /* 8-bit, f32 per-channel scale and offset         */

#define VFRB_ACC_F64   0  /* statistics accumulate in FP64 (reference)       */
#define VFRB_ACC_F32C  1  /* compensated (Neumaier) FP32                     */

typedef struct vfrb_search_desc {
    int32_t nchans;       /* channels (NCHANOUT)                             */
    double  tsamp;        /* s per row (tchan)                               */
    double  fch1;         /* MHz, centre of channel 0 (highest frequency)    */
    double  chanbw;       /* MHz, positive; channel j is fch1 - j*chanbw     */
    int32_t ndm;
    const float* dms;     /* host, ndm values                                */
    int32_t max_rows;     /* largest window (rows) ever passed               */
    int32_t storage;      /* VFRB_STORE_*                                    */
    int32_t tile_config;  /* -1 = default (the 4-channel group tile; for DM
                             lists too coarse for it the channel-pair tile,
                             then a per-channel tile), else an index returned
                             by vfrb_search_tile_count/vfrb_search_tile_name */
} vfrb_search_desc;

typedef struct vfrb_search_plan vfrb_search_plan;

/*
 * Replaces dedisp_create_plan + dedisp_set_dm_list + dedisp_get_dm_count +
 * dedisp_get_max_delay (PIPELINE:416-441) and the in-tree OFFSET table
 * (TIMESERIES:449-464).  The delay table is computed on the host in FP64
 * (vfrb_delay_table) and kept on the device.
 */
vfrb_status vfrb_search_plan_create(vfrb_search_plan** plan, const vfrb_search_desc* desc);
vfrb_status vfrb_search_plan_destroy(vfrb_search_plan* plan);
int32_t     vfrb_search_max_delay(const vfrb_search_plan* plan);
int32_t     vfrb_search_dm_count(const vfrb_search_plan* plan);
/* Copies the [ndm][nchans] int32 delay table to host memory. */
vfrb_status vfrb_search_get_offsets(const vfrb_search_plan* plan, int32_t* host_offsets);

/*
 * Loads one window into the plan: rows row0 .. row0+nrows-1 of a time-major
 * f32 filterbank ring (fb[(row mod ring_rows)*nchans + c]; ring_rows = 0 =
 * linear), applies subtractbpCAL when apply_bandpass != 0 (blocks of 128
 * rows aligned to row0, PIPELINE:1427-1474), and stores the residuals
 * channel-major in the plan's storage type.  This fuses PIPELINE:653-658
 * (tail save + bandpass) with the layout change the dedispersion wants.
 */
vfrb_status vfrb_search_load(vfrb_search_plan* plan, const float* fb,
                             int64_t ring_rows, int64_t row0, int32_t nrows,
                             int32_t apply_bandpass, void* stream);

/*
 * Replaces dedisp_execute(plan, nrows, fb, 32, out, 32, DEVICE_POINTERS)
 * (PIPELINE:671) with the arithmetic of the in-tree kernel `dedisperse`
 * (PIPELINE:1517-1537):
 *   out[dm][t] = (1/nchans) * sum_c FB[t + off[dm][c]][c],
 *   t < nrows - max_delay,
 * on the window last given to vfrb_search_load.  out is f32 [ndm][ld_out].
 */
vfrb_status vfrb_search_dedisperse(vfrb_search_plan* plan, int32_t nrows,
                                   float* out, int64_t ld_out, void* stream);

/*
 * Convenience: dedisp_execute on a caller-provided, already bandpassed,
 * time-major f32 filterbank (PIPELINE:671): load without bandpass, then
 * dedisperse.
 */
vfrb_status vfrb_search_execute(vfrb_search_plan* plan, int32_t nrows,
                                const float* fb_time_major, float* out,
                                int64_t ld_out, void* stream);

/*
 * dedisperseSTATS (TIMESERIES:1450-1481): mean over channels and the
 * channel-scatter term sqrt((S2 - S1*S1)/(N*(N-1))) with
 * S1 = sum(x-mean), S2 = sum((x-mean)^2) (the reference omits /N on S1*S1).
 * Works on f32 storage plans only.
 */
vfrb_status vfrb_search_dedisperse_stats(vfrb_search_plan* plan, int32_t nrows,
                                         float* out, float* sig, int64_t ld_out,
                                         void* stream);

/* Tile configurations of the dedispersion kernel (for the autotuner). */
int32_t     vfrb_search_tile_count(void);
const char* vfrb_search_tile_name(int32_t index);
vfrb_status vfrb_search_set_tile(vfrb_search_plan* plan, int32_t index);
int32_t     vfrb_search_get_tile(const vfrb_search_plan* plan);
/*
 * Autotuner (no reference counterpart): times every tile configuration
 * (samples per lane x DMs per wave x waves per workgroup) that fits the plan's
 * DM list on the window currently loaded, `iters` launches each, and keeps
 * the fastest.  ms (optional, host, vfrb_search_tile_count() floats) gets the
 * average milliseconds per launch, -1 where a configuration does not fit.
 * out is scratch with the vfrb_search_dedisperse shape.
 */
vfrb_status vfrb_search_autotune(vfrb_search_plan* plan, int32_t nrows, float* out,
                                 int64_t ld_out, int32_t iters, float* ms, void* stream);

/* ------------------------------------------------------------------------ */
/* Batched dedispersion (no reference counterpart: the reference runs one    */
/* dedisp_execute per window, PIPELINE:671)                                  */
/* ------------------------------------------------------------------------ */
/*
 * A batch dedisperses the loaded windows of many search plans (one per
 * antenna stream: own DM list, own window length) in ONE persistent launch
 * that pulls (window, DM block, time tile) work items from a list in device
 * memory.  Every member plan must have the batch's storage type; the batch
 * moves them to its tile configuration (vfrb_search_set_tile).  Each job's
 * output equals vfrb_search_dedisperse on that plan alone, bit for bit.
 */
typedef struct vfrb_batch vfrb_batch;

vfrb_status vfrb_batch_create(vfrb_batch** batch, int32_t storage, int32_t tile_config,
                              int32_t max_jobs, int32_t max_tiles);
vfrb_status vfrb_batch_destroy(vfrb_batch* batch);
/*
 * Sets the jobs of the next launch: job i dedisperses the first nrows[i] rows
 * of plans[i]'s loaded window into outs[i] (f32 [ndm][ld_outs[i]]).  The
 * work list is copied to the device on `stream`.  *ntiles (optional) receives
 * the number of work items.  Returns VFRB_ERR_CAPACITY beyond max_jobs or
 * max_tiles.
 */
vfrb_status vfrb_batch_set(vfrb_batch* batch, int32_t njobs, vfrb_search_plan* const* plans,
                           const int32_t* nrows, float* const* outs, const int64_t* ld_outs,
                           int32_t* ntiles, void* stream);
/* One eager persistent launch over the current jobs. */
vfrb_status vfrb_batch_dedisperse(vfrb_batch* batch, void* stream);
/*
 * hipGraph path: the first call captures the persistent launch on `stream`
 * (a stream from vfrb_stream_create, not the null stream) and instantiates
 * it; every call replays it.  The graph reads the work list from device
 * memory, so it stays valid across vfrb_batch_set calls; changing the tile
 * configuration drops it and the next call captures again.
 */
vfrb_status vfrb_batch_graph_launch(vfrb_batch* batch, void* stream);
/*
 * Autotuner over the batch: times every tile configuration that fits all
 * member plans on the current jobs (`iters` launches each), keeps the
 * fastest and rebuilds the work list for it.  ms as in vfrb_search_autotune.
 */
vfrb_status vfrb_batch_autotune(vfrb_batch* batch, int32_t iters, float* ms, void* stream);
int32_t     vfrb_batch_get_tile(const vfrb_batch* batch);
/* Persistent workgroups of one launch (occupancy x compute units). */
int32_t     vfrb_batch_grid(const vfrb_batch* batch);

/* ------------------------------------------------------------------------ */
/* Segment server (no reference counterpart: the reference processes one    */
/* VDIF stream, window by window, in a single host loop, PIPELINE:560-760)   */
/* ------------------------------------------------------------------------ */
/*
 * A request loop over many antenna streams.  Each stream keeps on the device
 * its raw filterbank ring (the max-delay overlap the next window needs, the
 * reference's tail save/restore of PIPELINE:653,696), the partial row of
 * payload bytes left over by the previous segment, its search plan (delay
 * table, window, storage scales) and its output buffers.  A segment of any
 * byte length is filterbanked onto the end of its stream; the window is the
 * last min(rows so far, max_delay + new rows) rows, bandpassed
 * (subtractbpCAL), dedispersed, and reduced to per-DM statistics (calcstats)
 * and a candidate list.
 *
 * batched = 1: continuous batching.  A scheduler thread admits waiting
 * segments (arrival order, one per stream at a time) while the device input
 * arena has room, up to max_batch per iteration, and runs an iteration on
 * whichever of two HIP streams is free, so a new iteration overlaps the one
 * in flight.  The dedispersion of all segments of an iteration is one
 * persistent vfrb_batch launch replayed from a captured hipGraph.
 * batched = 0: the plain path: one segment at a time, in arrival order, with
 * one dedispersion launch per segment.
 * Results are the same whichever segments share an iteration.
 */
typedef struct vfrb_server_desc {
    int32_t twonchan;      /* FFT length                                      */
    int32_t nave;          /* spectra per output row                          */
    int32_t nchans;        /* channels kept (NCHANOUT)                        */
    int32_t nskip;         /* first kept R2C bin (NSKIP)                      */
    double  fch1;          /* MHz, centre of channel 0                        */
    double  chanbw;        /* MHz                                             */
    double  tsamp;         /* s per row                                       */
    int32_t storage;       /* VFRB_STORE_* of the dedispersion window         */
    int32_t acc;           /* VFRB_ACC_* of the per-DM statistics             */
    float   threshold;     /* candidate SNR threshold (strict >)              */
    int32_t max_cand;      /* candidates kept per segment                     */
    int32_t max_streams;
    int64_t max_segment_bytes;  /* per polarisation                           */
    int64_t arena_bytes;   /* device input arena: admission limit             */
    int32_t max_batch;     /* segments per iteration (batched mode)           */
    int32_t batched;       /* 1 = continuous batching, 0 = plain FIFO         */
    int32_t tile_config;   /* dedispersion tile, -1 = default                 */
} vfrb_server_desc;

/* Caller-owned host buffers that receive one segment's results; any may be
 * NULL.  They must stay valid until the segment's result is polled. */
typedef struct vfrb_server_output {
    float*   peak;      /* [ndm] peak SNR per DM (calcstats, PIPELINE:1337-1370) */
    int32_t* peak_t;    /* [ndm] absolute output sample of that peak          */
    float*   ave;       /* [ndm]                                              */
    float*   sigma;     /* [ndm]                                              */
    int32_t* cand_dm;   /* [max_cand] candidates sorted by (dm, t)            */
    int32_t* cand_t;    /* [max_cand] absolute output sample                  */
    float*   cand_snr;  /* [max_cand]                                         */
} vfrb_server_output;

typedef struct vfrb_server_result {
    int64_t tag;
    int32_t stream_id;
    int32_t status;      /* VFRB_OK or this segment's error                  */
    int64_t rows;        /* filterbank rows the segment completed            */
    int64_t sample0;     /* absolute output sample of the window's t = 0     */
    int32_t nt;          /* samples searched; 0 while the stream holds no
                            more than max_delay + 1 rows                     */
    int32_t ncand;       /* candidates above threshold (may exceed max_cand) */
    int32_t best_dm;     /* DM index of the largest peak SNR, -1 if nt == 0  */
    int32_t best_t;      /* absolute sample of that peak                     */
    float   best_snr;
    int32_t batch_size;  /* segments in the same iteration                   */
    double  arrival_s;   /* server clock (s since create) at submit          */
    double  start_s;     /* iteration issued to the device                   */
    double  done_s;      /* results on the host                              */
} vfrb_server_result;

typedef struct vfrb_server_stats {
    int64_t submitted, completed, iterations, batched_segments;
    int64_t arena_peak_bytes;
    double  device_busy_s;  /* sum of iteration device times                 */
} vfrb_server_stats;

typedef struct vfrb_server vfrb_server;

vfrb_status vfrb_server_create(vfrb_server** server, const vfrb_server_desc* desc);
/* Waits for every submitted segment, then frees everything. */
vfrb_status vfrb_server_destroy(vfrb_server* server);
/* Adds a stream with its own DM list (host floats, ascending, last = largest).
 * Fails with VFRB_ERR_NOMEM when free device memory cannot hold its state. */
vfrb_status vfrb_server_open_stream(vfrb_server* server, const float* dms, int32_t ndm,
                                    int32_t* stream_id);
int32_t     vfrb_server_max_delay(const vfrb_server* server, int32_t stream_id);
/*
 * Queues nbytes of payload per polarisation (host memory, valid until the
 * result is polled) as the next segment of stream_id; returns at once.
 * VFRB_ERR_CAPACITY when nbytes exceeds max_segment_bytes.
 */
vfrb_status vfrb_server_submit(vfrb_server* server, int32_t stream_id, const uint8_t* pol0,
                               const uint8_t* pol1, int64_t nbytes, int64_t tag,
                               const vfrb_server_output* out);
/* Same, with the payload already in device memory (a NIC-to-GPU path):
 * no host-to-device copy; pol0/pol1 stay valid until the result is polled. */
vfrb_status vfrb_server_submit_device(vfrb_server* server, int32_t stream_id, const uint8_t* pol0,
                                      const uint8_t* pol1, int64_t nbytes, int64_t tag,
                                      const vfrb_server_output* out);
/*
 * Moves up to max_results finished results (completion order) into results;
 * waits up to wait_ms (< 0: until one is ready or nothing is outstanding) when
 * none is ready.  *n receives the count.
 */
vfrb_status vfrb_server_poll(vfrb_server* server, vfrb_server_result* results,
                             int32_t max_results, int32_t wait_ms, int32_t* n);
/* Blocks until every submitted segment has finished. */
vfrb_status vfrb_server_drain(vfrb_server* server);
vfrb_status vfrb_server_get_stats(const vfrb_server* server, vfrb_server_stats* stats);
/* Server clock: seconds since vfrb_server_create (steady clock). */
double      vfrb_server_now(const vfrb_server* server);
/*
 * Stream state (no reference counterpart: the reference keeps its overlap
 * tail powLEFTOVER in memory only, PIPELINE:463-464,653,696-701).  A stream's
 * state is the raw filterbank rows its next window starts with (the last
 * max_delay rows, f32), the payload bytes short of a row, and its row count.
 * save waits until the stream's submitted segments have finished, then
 * copies the state to host memory: *nbytes receives its size, buf NULL only
 * asks for it, VFRB_ERR_CAPACITY when cap is smaller.  restore puts a saved
 * state into a stream (of this or another server) opened with the same
 * filterbank shape, band, sample time and DM list (VFRB_ERR_INVALID
 * otherwise); its next segments then give the results the saved stream's
 * would.  Do not submit to the stream during either call.
 */
vfrb_status vfrb_server_save_stream(vfrb_server* server, int32_t stream_id, void* buf, int64_t cap,
                                    int64_t* nbytes);
vfrb_status vfrb_server_restore_stream(vfrb_server* server, int32_t stream_id, const void* buf,
                                       int64_t nbytes);

/* ------------------------------------------------------------------------ */
/* Per-DM statistics and candidates                                         */
/* ------------------------------------------------------------------------ */
/*
 * calcstats (PIPELINE:1337-1370): per DM, mean, two-pass corrected variance
 * (sum2 - sum*sum/N)/(N-1), ave=(float)mean, sigma=(float)sqrt(var),
 * SNR_t = (x_t - ave)/sigma in f32, peak = max(0, max SNR), frb = any SNR > 3.
 * acc = VFRB_ACC_F64 (reference) or VFRB_ACC_F32C (compensated FP32).
 * out is f32 [ndm][ld]; results are f32/int32 [ndm].  Any output pointer
 * except ave/sigma may be NULL.
 */
vfrb_status vfrb_dm_stats(const float* out, int64_t ld, int32_t ndm, int32_t nt,
                          int32_t acc, int32_t* frb, float* ave, float* sigma,
                          float* peak, int32_t* peak_t, void* stream);

/*
 * Candidate list (no reference counterpart; the reference only flags
 * SNR > 3 per DM): every (dm, t) with (x - ave[dm])/sigma[dm] > threshold
 * (strict, as PIPELINE:1363), appended through an atomic counter.  *count
 * (device int32, zeroed by this call on the stream) receives the total number
 * found, which may exceed max_cand (then only max_cand entries are stored,
 * in no particular order).
 */
vfrb_status vfrb_candidates(const float* out, int64_t ld, int32_t ndm, int32_t nt,
                            const float* ave, const float* sigma, float threshold,
                            int32_t* cand_dm, int32_t* cand_t, float* cand_snr,
                            int32_t max_cand, int32_t* count, void* stream);

/* ------------------------------------------------------------------------ */
/* VDIF (host)                                                              */
/* ------------------------------------------------------------------------ */
/* Decoded fields of the 32-byte VDIF header (vdifio.h:53-72). */
typedef struct vfrb_vdif_header {
    uint32_t seconds, legacymode, invalid;
    uint32_t frame, epoch, unassigned;
    uint32_t framelength8, nchan, version;
    uint32_t stationid, threadid, nbits, iscomplex;
    uint32_t eversion, extended1, extended2, extended3, extended4;
} vfrb_vdif_header;

vfrb_status vfrb_vdif_decode(const uint8_t* raw32, vfrb_vdif_header* out);
vfrb_status vfrb_vdif_encode(const vfrb_vdif_header* in, uint8_t* raw32);
/* getVDIFDMJD (external libvfdifio, called at PIPELINE:525): MJD (days, with
 * fraction) of the frame: epoch start + seconds + frame/framespersec. */
double      vfrb_vdif_dmjd(const uint8_t* raw32, int32_t framespersec);

/*
 * Frame sequencer (PIPELINE:829-1171): alternating pol-0/pol-1 frames, thread
 * alignment, pol frame-number agreement, start on the 10 Hz cal boundary
 * (frame % 2560 == 0), gap detection by frame number (mod 25600) with
 * zero-filled payloads, and partial-frame carry across bundles.
 * The log callback, when set, receives the reference's log lines.
 */
typedef struct vfrb_seq vfrb_seq;
typedef void (*vfrb_log_fn)(const char* line, void* user);

vfrb_status vfrb_seq_open_file(vfrb_seq** seq, const char* path);
vfrb_status vfrb_seq_open_memory(vfrb_seq** seq, const uint8_t* data, size_t bytes);
vfrb_status vfrb_seq_set_log(vfrb_seq* seq, vfrb_log_fn fn, void* user);
/* Gap rule.  0 (default): frame-number only, gap = (fnum - expected) mod
 * 25600 (PIPELINE:888-930).  1: seconds-aware, gap = fnum - expected +
 * (sec - expected_sec)*25600 and the pols must also agree on seconds at
 * init (TIMESERIES:893-937,1095-1140). */
vfrb_status vfrb_seq_set_seconds_check(vfrb_seq* seq, int32_t enable);
/* initDADA (TIMESERIES:767-826): discard npairs raw frame pairs before init. */
vfrb_status vfrb_seq_skip_pairs(vfrb_seq* seq, int64_t npairs);
/* initframe01 + initskip01 (PIPELINE:1044-1131).  Returns VFRB_ERR_EOF at
 * end of input and VFRB_ERR_FORMAT on an alignment failure. */
vfrb_status vfrb_seq_init(vfrb_seq* seq);
/* Header of the current pol-0 frame (after init: the first frame used). */
vfrb_status vfrb_seq_header(const vfrb_seq* seq, uint8_t* raw32);
/* filltimebufferHOST (PIPELINE:1135-1171): fills nloop bytes per pol.
 * Returns VFRB_ERR_EOF when the bundle cannot be completed (the reference
 * drops that bundle). */
vfrb_status vfrb_seq_fill(vfrb_seq* seq, int32_t nloop, uint8_t* out0, uint8_t* out1);
/* Frames read and frames zero-filled so far, per pol. */
vfrb_status vfrb_seq_counters(const vfrb_seq* seq, int64_t* frames_read, int64_t* frames_filled);
/* Input bytes the sequencer has read so far (whole 5032-byte frames).  After
 * vfrb_seq_init the pair it starts on (the cal-aligned thread 0/1 pair of
 * initframe01, PIPELINE:1058-1131) ends at this offset. */
vfrb_status vfrb_seq_consumed(const vfrb_seq* seq, int64_t* bytes);
vfrb_status vfrb_seq_close(vfrb_seq* seq);

/* ------------------------------------------------------------------------ */
/* Output formats (host)                                                    */
/* ------------------------------------------------------------------------ */
/* SIGPROC header of PIPELINE:529-540 (send_* of PIPELINE:1177-1207) into buf;
 * *len receives the byte count (needs <= 512 bytes). */
vfrb_status vfrb_sigproc_header(uint8_t* buf, int32_t cap, int32_t* len,
                                double fch1, double foff, int32_t nchans,
                                int32_t nbits, double tstart, double tsamp);
/* PRESTO .inf text of writeprestoheader (TIMESERIES:1483-1515). */
vfrb_status vfrb_presto_inf(char* buf, int32_t cap, const char* basename,
                            double mjdstart, int32_t ntimes, double tsamp,
                            float dm, double fchanlo, double bw, int32_t nchan,
                            double chanbw);

/* ------------------------------------------------------------------------ */
/* Synthetic VLITE voltages (host; counter-based RNG, any thread count)     */
/* ------------------------------------------------------------------------ */
#define VFRB_SYNTH_MAX_PULSES 64
#define VFRB_SYNTH_MAX_TONES  64

typedef struct vfrb_synth_spec {
    uint64_t seed;
    int32_t  nfft;        /* synthesis block length (= twonchan)             */
    double   mean, sigma; /* 128, 8                                           */
    double   f_top_mhz;   /* frequency of FFT bin 0 (384: inverted band)     */
    double   bw_mhz;      /* 64                                               */
    int32_t  npulse;
    double   pulse_dm[VFRB_SYNTH_MAX_PULSES];
    double   pulse_t0[VFRB_SYNTH_MAX_PULSES];    /* s, arrival at f_top      */
    double   pulse_fwhm[VFRB_SYNTH_MAX_PULSES];  /* s                         */
    double   pulse_amp[VFRB_SYNTH_MAX_PULSES];   /* fractional power excess   */
    int32_t  ntone;
    double   tone_bin[VFRB_SYNTH_MAX_TONES];     /* FFT bin (fractional ok)   */
    double   tone_amp[VFRB_SYNTH_MAX_TONES];     /* counts (peak)             */
    int32_t  nthreads;    /* 0 = all available                               */
} vfrb_synth_spec;

/* Samples [block0*nfft, (block0+nblocks)*nfft) of one polarisation. */
vfrb_status vfrb_synth_payload(const vfrb_synth_spec* spec, int32_t pol,
                               int64_t block0, int64_t nblocks, uint8_t* out);
/* Pulse amplitude giving a dedispersed peak SNR (SNR ~ A*sqrt(2*nave*nchan)). */
double      vfrb_synth_amp_for_snr(double snr, int32_t nave, int32_t nchan);

#ifdef __cplusplus
}
#endif
#endif /* VFRB_H */
