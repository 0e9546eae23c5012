/*
 * dspx.h -- C ABI of libdspx, the MI355X (gfx950) coherent-dedispersion
 * hot path.  Plain pointers and sizes only; no exceptions cross this ABI.
 * Every call returns a dspx status (0 = DSPX_OK) and, on failure, leaves a
 * message readable with dspx_last_error().
 *
 * All device pointers are HIP device memory on the context's device; all work
 * is enqueued on the context's stream and is asynchronous unless a function
 * says otherwise.  A context is NOT thread-safe (one per pipeline thread, like
 * the reference's one cudaStream per SingleThread, SingleThread.C:135-169).
 *
 * Reference interfaces replaced (paths relative to the dspsr source tree):
 *   dspx_unpack          GenericEightBitUnpacker::unpack_on_gpu
 *                        (Kernel/Classes/GenericEightBitUnpacker.C:92-131) and
 *                        SixteenBitUnpacker::unpack (SixteenBitUnpacker.C:23-78)
 *   dspx_dedisp_geometry PlasmaResponse::prepare (PlasmaResponse.C:91-119),
 *                        Response::set_optimal_ndat (Response.C:553-588),
 *                        Convolution::prepare (Convolution.C:316-455)
 *   dspx_conv_*          Convolution::Engine prepare / copy_response / perform
 *                        (Convolution.C:446-451, 656-667, 690-697;
 *                        ConvolutionCUDA.cu:174-216, 219-284, 567-770)
 *   dspx_detect          Detection::Engine square_law / polarimetry
 *                        (Detection.C:237-243, 483-490; DetectionCUDA.cu:143,294)
 *   dspx_fold_*          Fold::Engine set_bins / fold / get_bin_hits
 *                        (Signal/Pulsar/dsp/Fold.h:262-345; Fold.C:760-887)
 *   dspx_sk_*            SpectralKurtosis::Engine compute / detect_ft / mask
 *                        (SpectralKurtosis.C:699-703, 1130-1134, 1518-1524)
 *                        and detect_tscr / get_limits (:989-1103)
 *   dspx_timediv_*       TimeDivide::set_bounds / set_boundaries
 *                        (Signal/Pulsar/TimeDivide.C:36-506, Subint.h:253-330)
 *   dspx_ltf_*           the LoadToFold operation chain Unpack -> Convolution
 *                        -> Detection -> Fold (LoadToFold1.C:109-258,
 *                        SingleThread.C:662-769) fused into one kernel or
 *                        three, with SpectralKurtosis excision (:265-377).
 */
#ifndef DSPX_H
#define DSPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DSPX_OK = 0,
  DSPX_E_PARAM = 1,       /* reference: Error(InvalidParam, ...) */
  DSPX_E_STATE = 2,       /* reference: Error(InvalidState, ...) */
  DSPX_E_HIP = 3,         /* reference: Error(FailedCall, ...) / CUFFTError */
  DSPX_E_UNSUPPORTED = 4, /* shape outside what the kernels were built for */
  DSPX_E_INTERNAL = 5
};

typedef struct dspx_ctx dspx_ctx;

/* Context: device, stream (nullable: the library creates one) and the
   twiddle tables shared by every engine on it. */
int dspx_ctx_create(int device, void* hip_stream, dspx_ctx** out);
void dspx_ctx_destroy(dspx_ctx* ctx);
const char* dspx_last_error(const dspx_ctx* ctx);
int dspx_ctx_synchronize(dspx_ctx* ctx);
void* dspx_ctx_stream(dspx_ctx* ctx);
/* Library version string, e.g. "dspx 0.1 gfx950". */
const char* dspx_version(void);

/* ---------------------------------------------------------------- geometry */

typedef struct {
  double centre_frequency;   /* MHz */
  double bandwidth;          /* MHz; negative for an inverted band */
  double dispersion_measure; /* pc cm^-3 */
  uint32_t nchan;
  uint32_t dc_centred;       /* Observation::get_dc_centred */
  uint32_t nfft;             /* 0: reference rule; else user-set resolution */
  uint32_t interchan;        /* 1: inter-channel delays removed (dspsr -K): the chirp
                                carries each channel's fractional sample delay
                                (Dedispersion.C:142-178) and dspx_delay_* applies the
                                integer delays */
} dspx_dedisp_desc;

typedef struct {
  uint32_t impulse_neg;   /* nfilt_neg: samples lost at the end of each FFT */
  uint32_t impulse_pos;   /* nfilt_pos: samples lost at the start */
  uint32_t nfft;
  uint32_t nsamp_overlap; /* impulse_neg + impulse_pos */
  uint32_t nsamp_step;    /* nfft - nsamp_overlap */
  uint32_t nunsupported;  /* leading channels whose smear exceeds the limit */
} dspx_geometry;

/* Host-only.  Complex (Analytic) input. */
int dspx_dedisp_geometry(const dspx_dedisp_desc* desc, dspx_geometry* out);

/* Host-only.  The inter-channel dispersion delays of SampleDelay with the
   Dedispersion::SampleDelay function (SampleDelay.C:56-160,
   DedispersionSampleDelay.C:25-129; one span, every pol alike) at the
   channel sample rate |bw|/nchan MHz:
     delay[c]   = floor(s_c + 0.5), frac[c] = s_c - delay[c], with
     s_c = (DM / 2.41e-4) * (1/f0^2 - 1/f_c^2) * rate   (f0 band centre,
     f_c channel centre),
     applied[c] = zero_delay + centre - delay[c] (input sample i + applied[c]
     becomes output sample i), total = max_c applied[c];
   zero_delay shifts the output start time (SampleDelay.C:221).  Any output
   pointer may be null. */
int dspx_sample_delays(const dspx_dedisp_desc* desc, int64_t* delay, double* frac, int64_t* applied,
                       int64_t* zero_delay, uint64_t* total_delay);

/* ------------------------------------------------------------- DADA source */

/* Host-only.  ascii_header_get (Kernel/Classes/ascii_header.c): keyword key
   counts where it starts the header, or follows a newline (or a backslash)
   and is followed by a space or tab; its value is the first
   whitespace-delimited token after it (sscanf "%s").  Copies the token,
   NUL-terminated and truncated to cap bytes, and returns 1; -1 when the key
   is absent or has no value. */
int dspx_ascii_header_get(const char* header, const char* key, char* value, uint32_t cap);

typedef struct {
  double centre_frequency; /* FREQ, MHz */
  double bandwidth;        /* BW, MHz (negative: inverted band) */
  double tsamp;            /* TSAMP, microseconds */
  uint64_t hdr_size;       /* HDR_SIZE, bytes of header before the data; 0 if absent */
  uint64_t obs_offset;     /* OBS_OFFSET (or OFFSET), bytes from the start of the observation */
  uint64_t ndat;           /* NDAT; 0 if absent */
  int64_t picoseconds;     /* PICOSECONDS; 0 if absent */
  uint32_t nchan;          /* NCHAN; 1 if absent */
  uint32_t npol;           /* NPOL */
  uint32_t nbit;           /* NBIT */
  uint32_t ndim;           /* NDIM; 1 if absent (1 Nyquist, 2 Analytic, 4 Coherence) */
  uint32_t resolution;     /* RESOLUTION in bytes; 1 if absent */
  uint32_t reserved;
  char utc_start[32];      /* UTC_START as written (YYYY-MM-DD-hh:mm:ss) */
  char source[64];
  char telescope[64];
  char instrument[64];     /* empty if absent */
} dspx_dada_info;

/* Host-only.  The header fields that ASCIIObservation::load reads for this
   path (ASCIIObservation.C:24-45, 100-420): returns DSPX_E_PARAM when a
   required key (TELESCOPE, SOURCE, FREQ, BW, NPOL, NBIT, TSAMP, UTC_START,
   OBS_OFFSET) is absent or does not parse, naming it in missing (>= 32 bytes,
   may be null), and DSPX_E_STATE for NDIM other than 1, 2 or 4. */
int dspx_dada_parse(const char* header, dspx_dada_info* out, char* missing);

/* ------------------------------------------------------------- SampleDelay */

/* SampleDelay::Engine (SampleDelayCUDA.cu:24-120): set_delays once, then
   retard per block. */
typedef struct dspx_delay dspx_delay;
/* applied: host [nchan * npol] sample offsets (ichan * npol + ipol), >= 0. */
int dspx_delay_create(dspx_ctx* ctx, uint32_t nchan, uint32_t npol, const int64_t* applied, dspx_delay** out);
void dspx_delay_destroy(dspx_delay* delay);
/* out[c][p][i] = in[c][p][i + applied[c*npol + p]] for i < ndat_out, ndim
   floats per sample (FPT; strides in floats).  in != out. */
int dspx_delay_retard(dspx_delay* delay, const float* d_in, uint64_t in_chan_stride, uint64_t in_pol_stride,
                      uint32_t ndim, uint64_t ndat_out, float* d_out, uint64_t out_chan_stride,
                      uint64_t out_pol_stride);

/* ------------------------------------------------------------------ unpack */

/* Raw TFP [ndat][nchan][npol][2] (8-bit: (v + 0.5) * scale, offset binary
   (v - 127.5) * scale; 16-bit: float(v) + 0.5, scale ignored, offset binary
   XOR 0x8000) -> FPT float complex out[ichan*chan_stride + ipol*pol_stride
   + 2*idat (+1)], strides in floats. */
int dspx_unpack(dspx_ctx* ctx, const void* d_raw, uint32_t nbit, uint32_t offset_binary,
                uint32_t nchan, uint32_t npol, uint64_t ndat, float scale, float* d_out,
                uint64_t out_chan_stride, uint64_t out_pol_stride);

/* ------------------------------------------------------------- convolution */

enum {
  DSPX_CHIRP_TABLE = 0,  /* response copied from the host (copy_response) */
  DSPX_CHIRP_PARITY = 1, /* on-chip FP64 phase rounded to float, as the
                            reference stores it (Dedispersion.C:104-111) */
  DSPX_CHIRP_TRUTH = 2   /* on-chip FP64 phase, no float rounding */
};

enum {
  DSPX_FORM_FUSED = 0,    /* hand-written: the one-pass kernel where the transform
                             (2^7..2^12 points) and the fold histogram fit the LDS
                             (LoadToFold plans), else the three stages */
  DSPX_FORM_PLAIN = 1,    /* hipFFT + separate multiply / discard kernels */
  DSPX_FORM_TWOLEVEL = 2, /* the three-stage (four-step) form, 2^12..2^22 points */
  DSPX_FORM_ONEPASS = 3   /* the one-pass form, 2^7..2^13 points (LoadToFold plans) */
};

typedef struct dspx_conv dspx_conv;

typedef struct {
  dspx_dedisp_desc dedisp; /* band and DM for the on-chip chirp */
  uint32_t nchan;
  uint32_t npol;
  uint32_t nfft;
  uint32_t nfilt_pos;
  uint32_t nfilt_neg;
  uint32_t chirp_mode;
  uint32_t form;
  uint32_t nyquist;        /* 0: Analytic (complex) input; 1: real-sampled
                              (Signal::Nyquist) input, fused forms only */
} dspx_conv_desc;

int dspx_conv_create(dspx_ctx* ctx, const dspx_conv_desc* desc, dspx_conv** out);
void dspx_conv_destroy(dspx_conv* conv);
/* Engine::copy_response: [nchan][nfft] complex float in FFT-bin order, the
   normaliser already multiplied in.  Required for DSPX_CHIRP_TABLE. */
int dspx_conv_set_response(dspx_conv* conv, const float* host_response);
/* Engine::perform: FPT complex input (ndat samples per chan/pol), npart
   overlapping FFTs of nfft with step nfft - nfilt_pos - nfilt_neg; output
   npart*step samples.  d_zero_dm (nullable) receives the zero-DM copy.
   Strides in floats.  Real-sampled input (nyquist = 1, Convolution.C:385-413,
   597-608, 700-790, ConvolutionCUDA.cu:776-944): ndat real samples per
   chan/pol; part p is the 2*nfft real samples from 2*p*step, whose 2*nfft-
   point real transform's bins 0..nfft-1 take the response (in frequency
   order, not swapped; the on-chip chirp zeroes the first channel's DC, as
   Response.C:418-420 does) before the nfft-point backward transform; the
   output is complex, npart*step samples at half the input rate; no zero-DM
   copy. */
int dspx_conv_perform(dspx_conv* conv, const float* d_in, uint64_t in_chan_stride,
                      uint64_t in_pol_stride, uint64_t ndat, float* d_out,
                      uint64_t out_chan_stride, uint64_t out_pol_stride, float* d_zero_dm,
                      uint32_t npart);

/* --------------------------------------------------------------- filterbank */
/* Filterbank::Engine (Filterbank.C:245-250, 463-474, 549-561;
   FilterbankEngineCPU.C:184-370): nchan input channels each split into nsub
   sub-bands of fres frequency bins, convolved with the response on the way:
   the nsub*fres-point forward transform of the window at ipart*nsub*nkeep,
   the response product, the fres-point backward transform of each
   sub-band's slice, the kept samples [nfilt_pos, nfilt_pos + nkeep) (nkeep =
   fres - nfilt_pos - nfilt_neg).  Three forms, chosen by window shape:
   one kernel per window (1024-4096 points, fres 32-512); the four-step form
   (2^13-2^22 points with fres >= 8*N1, N1 = 16 below 2^16 points and
   nfft/4096 from there: forward column transform, rows with their sub-band
   backward transforms, backward column transform into the kept samples);
   batched FFT-library transforms otherwise (or with the environment
   DSPX_FB_FORM=library). */
typedef struct dspx_fb dspx_fb;

typedef struct {
  uint32_t nchan;     /* input channels */
  uint32_t npol;
  uint32_t nsub;      /* output channels per input channel (nchan_subband) */
  uint32_t fres;      /* frequency resolution: bins per output channel */
  uint32_t nfilt_pos; /* the response's impulse_pos / impulse_neg */
  uint32_t nfilt_neg;
  uint32_t reserved[2];
} dspx_fb_desc;

int dspx_fb_create(dspx_ctx* ctx, const dspx_fb_desc* desc, dspx_fb** out);
void dspx_fb_destroy(dspx_fb* fb);
/* The response [nchan][nsub*fres] complex float, in the order the
   reference's Response holds it after Response::swap_as_needed
   (Response.C:325-400) and with the normaliser multiplied in
   (Filterbank.C:176-185).  Synchronous. */
int dspx_fb_set_response(dspx_fb* fb, const float* host_response);
/* Parts and output samples per output channel for ndat input samples
   (Filterbank::resize_output, Filterbank.C:410-440). */
int dspx_fb_geometry(const dspx_fb* fb, uint64_t ndat, uint64_t* npart, uint64_t* ndat_out);
/* FPT complex input [nchan][npol][ndat] (strides in floats) -> FPT complex
   output [nchan*nsub][npol][npart*nkeep]. */
int dspx_fb_perform(dspx_fb* fb, const float* d_in, uint64_t in_chan_stride, uint64_t in_pol_stride,
                    uint64_t ndat, float* d_out, uint64_t out_chan_stride, uint64_t out_pol_stride);
/* dspx_fb_perform with the zero-DM output as well: d_zdm (same layout and
   strides as d_out) receives the sub-bands of each window's spectrum
   transformed back without the response (no zero-DM response set, as in
   LoadToFold::construct for a filterbank that convolves during,
   LoadToFold1.C:291-303), for the spectral kurtosis of undispersed data.
   Replaces Filterbank::Engine::perform(in, out, zero_DM_out, npart,
   in_step, out_step) (Filterbank.C:549-555; FilterbankEngineCPU.C:266-300,
   FilterbankEngineCUDA.cu:329-344). */
int dspx_fb_perform_zdm(dspx_fb* fb, const float* d_in, uint64_t in_chan_stride, uint64_t in_pol_stride,
                        uint64_t ndat, float* d_out, uint64_t out_chan_stride, uint64_t out_pol_stride,
                        float* d_zdm);
/* The number of FFT-library batches the last dspx_fb_perform ran (0 when it
   ran the one-kernel or the four-step form); for checking the batch loop. */
int dspx_fb_batches(const dspx_fb* fb, uint32_t* nbatch);

/* --------------------------------------------------------------- detection */

enum { DSPX_INTENSITY = 0, DSPX_PPQQ = 1, DSPX_COHERENCE = 2, DSPX_STOKES = 3 };

/* FPT complex input -> FPT detected output.  Intensity: out[ichan*chan_stride
   + idat]; PPQQ: ipol*pol_stride added; Coherence/Stokes with ndim 1, 2 or 4
   lay the four products out as Detection::get_result_pointers does
   (Detection.C:578-624): value i of sample idat at
   (i/ndim)*pol_stride + idat*ndim + i%ndim. */
int dspx_detect(dspx_ctx* ctx, uint32_t state, uint32_t ndim, const float* d_in,
                uint64_t in_chan_stride, uint64_t in_pol_stride, uint32_t nchan, uint32_t npol,
                uint64_t ndat, float* d_out, uint64_t out_chan_stride, uint64_t out_pol_stride);

/* -------------------------------------------------------------------- fold */

/* One Fold::fold call: samples [idat_start, idat_start + ndat) of the block
   go to sub-integration isub, starting from phase phi = get_phi(start time of
   the midpoint of sample idat_start) (Fold.C:686-693). */
typedef struct {
  uint64_t idat_start;
  uint64_t ndat;
  double phi;
  uint32_t isub;
  uint32_t reserved;
} dspx_fold_chunk;

typedef struct dspx_fold dspx_fold;

/* Builds the bin plan on the device for a block of ndat samples.  The bins
   reproduce the reference recurrence phi -= floor(phi); ibin =
   unsigned(phi*nbin); phi += pps (Fold.C:790-838) bit for bit. */
int dspx_fold_create(dspx_ctx* ctx, uint64_t ndat, double phase_per_sample, uint32_t nbin,
                     const dspx_fold_chunk* chunks, uint32_t nchunk, dspx_fold** out);
void dspx_fold_destroy(dspx_fold* fold);
/* Options of Fold::fold beyond the phase model (Fold.C:662-740):
   WeightedTimeSeries weights (sample idat of the block has weight
   weights[(idat + weight_idat) / ndatperweight]; a zero weight marks the
   sample not folded, Fold.C:792-810, 826-827), and zeroed data
   (input->get_zeroed_data(): the weights mark nothing and the hits are
   counted by dspx_fold_execute_zeroed, Fold.C:893-930).  weights is read
   during the call only. */
typedef struct {
  const uint16_t* weights;   /* host; NULL with ndatperweight = 0 */
  uint64_t nweights;
  uint32_t ndatperweight;
  uint32_t weight_idat;
  uint32_t zeroed_samples;
  uint32_t reserved;
} dspx_fold_opts;
int dspx_fold_create_opts(dspx_ctx* ctx, uint64_t ndat, double phase_per_sample, uint32_t nbin,
                          const dspx_fold_chunk* chunks, uint32_t nchunk, const dspx_fold_opts* opts,
                          dspx_fold** out);
/* Rebuilds the plan of an existing fold for the next block (the per-block
   Fold::fold call), reusing its device buffers; nothing waits for the
   stream. */
int dspx_fold_replan(dspx_fold* fold, uint64_t ndat, double phase_per_sample, uint32_t nbin,
                     const dspx_fold_chunk* chunks, uint32_t nchunk, const dspx_fold_opts* opts);
/* Host-side count of the samples the plan folds (Fold::fold's ndat_folded,
   Fold.C:831-836): the chunks' samples less those of zero weight; 0 with
   zeroed data, whose count is made on the device by dspx_fold_execute_zeroed. */
int dspx_fold_ndat_folded(const dspx_fold* fold, uint64_t* ndat_folded);
/* Per-sample plan: (isub << 16) | ibin, or 0xffffffff for samples outside
   every chunk.  d_out holds ndat uint32. */
int dspx_fold_binplan(dspx_fold* fold, uint32_t* d_out);
/* hits[isub][ibin] += number of samples folded into ibin (Fold.C:827-836). */
int dspx_fold_hits(dspx_fold* fold, uint32_t* d_hits);
/* Host-only evaluation of the same plan for one chunk starting at sample 0
   (bins[j] for j < ndat), and the number of arithmetic runs it uses; for
   checking the plan against the serial recurrence without a device. */
int dspx_fold_binplan_host(double phi, double phase_per_sample, uint32_t nbin, uint64_t ndat,
                           uint32_t* bins);
int dspx_fold_nsegments_host(double phi, double phase_per_sample, uint64_t ndat, uint64_t* nseg);
/* Engine::fold: detected FPT input (value idim of sample idat of (ichan, ipol)
   at ichan*chan_stride + ipol*pol_stride + idat*ndim + idim) is added into
   prof[isub][ichan][ipol][ibin*ndim + idim] (accumulated). */
int dspx_fold_execute(dspx_fold* fold, const float* d_in, uint64_t in_chan_stride,
                      uint64_t in_pol_stride, uint32_t nchan, uint32_t npol, uint32_t ndim,
                      float* d_prof);
/* Engine::fold on zeroed data (Fold.C:893-930; FoldCUDA.cu:424-583): as
   dspx_fold_execute, and hits[isub][hits_nchan][nbin] += the samples of pol 0
   whose value 0 is non-zero (hits_nchan = nchan: per channel; 1: all
   channels into one row), their number added into *d_ndat_folded (device). */
int dspx_fold_execute_zeroed(dspx_fold* fold, const float* d_in, uint64_t in_chan_stride,
                             uint64_t in_pol_stride, uint32_t nchan, uint32_t npol, uint32_t ndim,
                             float* d_prof, uint32_t* d_hits, uint32_t hits_nchan,
                             unsigned long long* d_ndat_folded);

/* The fold without allocation: dspx_fold_create + dspx_fold_hits +
   dspx_fold_execute in one call that allocates nothing (no hipMalloc,
   hipFree or allocating library routine).  It has no counterpart in the
   reference; it is for callers that cannot allocate device memory (tight
   serving loops, harnesses that forbid it).  The plan's segment table lives
   in a fixed device array of 32768 segments (1.25 MiB per GPU, held until
   the process ends); a plan needing more is refused with
   DSPX_E_UNSUPPORTED.  The bins are evaluated from the segments in the
   kernels (no per-sample plan).  chunks must be sorted by idat_start.  The
   call waits for the context's stream once (the table's upload, ordered
   after earlier work); the kernels then run asynchronously.  Every call on a
   GPU shares that GPU's table: calls that overlap on one GPU (several
   threads or contexts) race.  d_hits (nullable): hits[isub][nbin] +=;
   d_in and d_prof (both or neither): as dspx_fold_execute. */
int dspx_fold_static(dspx_ctx* ctx, uint64_t ndat, double phase_per_sample, uint32_t nbin,
                     const dspx_fold_chunk* chunks, uint32_t nchunk, const float* d_in,
                     uint64_t in_chan_stride, uint64_t in_pol_stride, uint32_t nchan, uint32_t npol,
                     uint32_t ndim, float* d_prof, uint32_t* d_hits);

/* ---------------------------------------------------- sub-integrations */

/* dsp::TimeDivide (Signal/Pulsar/TimeDivide.C:36-506) with divisions of a
   fixed number of seconds, or of turns of a constant folding period (the
   branch Subint::prepare takes without a predictor, Subint.h:217-240):
   which samples of each block belong to which sub-integration, continued
   across blocks.  Host computation; no device work. */
typedef struct {
  int64_t day;  /* MJD: integer day */
  int64_t sec;  /* seconds of the day */
  double frac;  /* fraction of a second */
} dspx_mjd;

typedef struct dspx_timediv dspx_timediv;

/* The samples [idat_start, idat_start + ndat) of a block in division
   `division` (counted from the start time) of division_ndat samples, and
   TimeDivide's flags: the division is new (bounds started a new one), its
   end is in this block, the block has samples for a later division. */
typedef struct {
  uint64_t idat_start;
  uint64_t ndat;
  uint64_t division;
  uint64_t division_ndat;
  int32_t new_division;
  int32_t end_reached;
  int32_t in_next;
  int32_t is_valid;
} dspx_timediv_bounds;

/* set_start_time / set_seconds / set_turns / set_period / set_reference_epoch
   (TimeDivide.C:36-104): exactly one of seconds and turns positive; turns
   need period (s).  Without a reference epoch, a whole number of seconds
   per division rounds the start down to a multiple of it within the day
   (TimeDivide.C:55-65).  reference_epoch may be NULL. */
int dspx_timediv_create(const dspx_mjd* start_time, double seconds, double turns, double period,
                        const dspx_mjd* reference_epoch, dspx_timediv** out);
void dspx_timediv_destroy(dspx_timediv* td);
/* One TimeDivide::set_bounds call (TimeDivide.C:124-338) for a block of
   input_ndat samples at `rate` (Hz) starting at input_start; division edges
   are rounded to the block's samples (set_boundaries, :467-506). */
int dspx_timediv_set_bounds(dspx_timediv* td, const dspx_mjd* input_start, double rate, uint64_t input_ndat,
                            dspx_timediv_bounds* out);
/* The set_bounds calls Subint::transformation makes for one block (until
   in_next is clear, Subint.h:253-330): the block's valid bounds, in order,
   into out[max]; *n of them.  DSPX_E_PARAM when more than max. */
int dspx_timediv_block(dspx_timediv* td, const dspx_mjd* input_start, double rate, uint64_t input_ndat,
                       dspx_timediv_bounds* out, uint32_t max, uint32_t* n);

/* ------------------------------------------------------- spectral kurtosis */
/* One resolution of dsp::SpectralKurtosis: blocks of M samples starting
   every M / noverlap samples (Resolution::prepare, SpectralKurtosis.C:
   263-281), the averaging detectors over the included channels
   (set_channel_range, :848-851; Resolution::get_channels, :342-360).  Sums
   [npart][nchan][npol][2] float (S1, S2), counts [npart][nchan] uint32, mask
   [npart][nchan] bytes (1 = zapped), as the reference's sums
   WeightedTimeSeries and zapmask BitSeries.  Coarser resolutions (adaptive
   SK, the reference's CPU path; its Engine path runs the first only,
   SpectralKurtosis.C:1131-1135) are added with dspx_sk_add_resolution and
   run by the _res entry points below. */
typedef struct dspx_sk dspx_sk;

typedef struct {
  uint32_t nchan;
  uint32_t npol;          /* pols of the sums (1 or 2) */
  uint32_t M;             /* samples per block */
  uint32_t Nd;            /* integrated_Nd: 1 for complex (Analytic) input */
  uint32_t omit_outliers; /* SpectralKurtosis::set_omit_outliers */
  uint32_t noverlap;      /* SpectralKurtosis::set_noverlap; 0 or 1: blocks do not overlap */
} dspx_sk_desc;

enum { DSPX_SK_SKFB = 1, DSPX_SK_FSCR = 2, DSPX_SK_TSCR = 4 };

int dspx_sk_create(dspx_ctx* ctx, const dspx_sk_desc* desc, dspx_sk** out);
void dspx_sk_destroy(dspx_sk* sk);
/* Thresholds by block count 0..M (host arrays of M+1 floats: the SKLimits
   lower / upper / symmetric thresholds of each count, as the reference's
   SpectralKurtosis::get_limits caches them, SpectralKurtosis.C:989-1009) and
   the per-comparison sigma of the closed form that detect_tscr uses for
   totals >= 32768 (SKLimits.C:152-160).  Synchronous. */
int dspx_sk_set_limits(dspx_sk* sk, const float* lower, const float* upper, const float* symmetric,
                       double pc_std_devs);
/* The blocks of a series of ndat samples (Resolution::prepare,
   SpectralKurtosis.C:263-281): npart blocks, and the ndat_out samples of the
   masked output (dspx_sk_mask).  Host only. */
int dspx_sk_blocks(const dspx_sk* sk, uint64_t ndat, uint64_t* npart, uint64_t* ndat_out);
/* The channels detect_fscr averages and detect_tscr tests: include[nchan]
   host bytes, non-zero where included (the include / exclude ranges of
   Resolution::get_channels resolved per channel); NULL includes every
   channel.  detect_skfb tests every channel, as the reference's does.
   Synchronous. */
int dspx_sk_set_channels(dspx_sk* sk, const uint8_t* include);
/* Engine::compute (SpectralKurtosis.C:699-703, 674-798) on an FPT complex
   float TimeSeries: blocks [0, npart), block b the M samples from b * M /
   noverlap.  zeroed_data: zero samples are not counted
   (input->get_zeroed_data()). */
int dspx_sk_compute(dspx_sk* sk, const float* d_in, uint64_t in_chan_stride, uint64_t in_pol_stride,
                    uint64_t npart, int zeroed_data, float* d_sums, uint32_t* d_counts);
/* The same from raw TFP 8/16-bit samples (unpacked as dspx_unpack does)
   starting at input sample t0: the zero-DM input of a convolution whose
   output sample 0 is input sample t0 = nfilt_pos (Convolution.C:792-799). */
int dspx_sk_compute_raw(dspx_sk* sk, const void* d_raw, uint32_t nbit, uint32_t offset_binary, float scale,
                        uint64_t t0, uint64_t npart, float* d_sums, uint32_t* d_counts);
/* SpectralKurtosis::detect (SpectralKurtosis.C:830-909): the mask is reset,
   then detect_skfb (DSPX_SK_SKFB, :1105-1222), detect_fscr (DSPX_SK_FSCR,
   :1312-1462) and detect_tscr (DSPX_SK_TSCR, :1014-1103) as flagged.  With
   omit_outliers, detect_skfb zeroes the sums and counts of the blocks it
   zaps.  detect_tscr tests every channel's total count with the thresholds
   of SKLimits(total) (get_limits, SpectralKurtosis.C:989-1012, 1045): those
   of dspx_sk_set_limits up to M, those of dspx_sk_set_tscr_limits above it,
   else from 32768 on the closed form of SKLimits.C:152-160.  With
   DSPX_SK_TSCR the call waits for the stream and fails with DSPX_E_STATE
   (the mask then misses those channels' tests) when a total in (M, 32768)
   has no thresholds: dspx_sk_tscr_totals names the totals to evaluate. */
int dspx_sk_detect(dspx_sk* sk, float* d_sums, uint32_t* d_counts, uint64_t npart, uint32_t flags,
                   uint8_t* d_mask);
/* The per-channel totals of the counts that detect_tscr will test
   (SpectralKurtosis.C:1036-1040): h_totals[nchan], host memory.  Called
   after a dspx_sk_detect without DSPX_SK_TSCR, it sees the counts that
   detect_skfb left (omit_outliers zeroes those of zapped blocks).
   Synchronous. */
int dspx_sk_tscr_totals(dspx_sk* sk, const uint32_t* d_counts, uint64_t npart, uint64_t* h_totals);
/* detect_tscr's thresholds of totals above M: n ascending totals with the
   lower and upper thresholds of SKLimits(total) (Nd and ntest = npol as in
   get_limits, SpectralKurtosis.C:989-1009).  Replaces the previous table.
   Synchronous. */
int dspx_sk_set_tscr_limits(dspx_sk* sk, uint32_t n, const uint64_t* totals, const float* lower,
                            const float* upper);
/* detect_tscr alone (SpectralKurtosis.C:1014-1103), OR-ed into the mask of a
   preceding dspx_sk_detect: the second half of a detect whose totals were
   read with dspx_sk_tscr_totals.  Waits for the stream; DSPX_E_STATE as
   dspx_sk_detect. */
int dspx_sk_detect_tscr(dspx_sk* sk, const float* d_sums, const uint32_t* d_counts, uint64_t npart,
                        uint8_t* d_mask);
/* SpectralKurtosis::mask (SpectralKurtosis.C:1465-1566) with the npart
   blocks of the sums: the output's ndat_out samples (dspx_sk_blocks) are
   input samples M - M/noverlap on, each zeroed in every pol where the last
   block written over it is flagged (noverlap 1: out = in with block b of
   channel c zeroed where mask[b][c]; in may then equal out). */
int dspx_sk_mask(dspx_sk* sk, const uint8_t* d_mask, const float* d_in, uint64_t in_chan_stride,
                 uint64_t in_pol_stride, float* d_out, uint64_t out_chan_stride, uint64_t out_pol_stride,
                 uint64_t npart);

/* Several resolutions (SpectralKurtosis::Resolution, SpectralKurtosis.C:
   101-211, 251-294, 360-380).  The engine's own M and noverlap are the first
   (finest); each added resolution has M a multiple of the previous one's
   and a block offset M/noverlap a multiple of the previous one's
   (Resolution::compatible), added in increasing M, with its thresholds by
   count 0..M (as dspx_sk_set_limits) and its included channels (include
   [nchan] host bytes, NULL: every channel).  Synchronous. */
int dspx_sk_add_resolution(dspx_sk* sk, uint32_t M, uint32_t noverlap, const float* lower, const float* upper,
                           const float* symmetric, const uint8_t* include);
/* Removes the added resolutions. */
int dspx_sk_clear_resolutions(dspx_sk* sk);
/* The blocks of a series of ndat samples with the resolutions: npart of the
   first (the mask's rows) and the ndat_out samples of the masked output
   (those of the last resolution, SpectralKurtosis::reserve, :454-480).  Host
   only. */
int dspx_sk_blocks_res(const dspx_sk* sk, uint64_t ndat, uint64_t* npart, uint64_t* ndat_out);
/* SpectralKurtosis::detect with every resolution (:849-905) over the first
   resolution's sums of a series of ndat samples (dspx_sk_compute over
   dspx_sk_blocks_res's npart blocks): for each resolution in turn, the sums
   time-scrunched from the previous one (tscrunch_sums, :917-972: block b of
   the coarser is the sum of blocks (b + i) * offset, i < nsum, of the finer,
   nsum the ratio of their M and offset that of their block offsets; the
   reference asserts that the last block read exists, DSPX_E_STATE here),
   then detect_skfb and detect_fscr as flagged, their zaps flagging the
   first resolution's rows that the block spans (:1121-1131, 1355-1364);
   finally detect_tscr (DSPX_SK_TSCR) over the last resolution's blocks and
   thresholds with the first resolution's channels, flagging the mask's first
   rows, as many as the last resolution has blocks (:1087-1093).  The
   coarser sums are the engine's own: d_sums and d_counts keep the first
   resolution's (zeroed where omit_outliers zaps).  With DSPX_SK_TSCR the
   call waits for the stream; DSPX_E_STATE as dspx_sk_detect. */
int dspx_sk_detect_res(dspx_sk* sk, float* d_sums, uint32_t* d_counts, uint64_t ndat, uint32_t flags,
                       uint8_t* d_mask);
/* After a dspx_sk_detect_res without DSPX_SK_TSCR: the per-channel totals
   detect_tscr will test (h_totals[nchan], host), and detect_tscr itself
   OR-ed into d_mask (waits for the stream; DSPX_E_STATE as dspx_sk_detect).
   They read the last resolution's sums that call left. */
int dspx_sk_res_totals(dspx_sk* sk, uint64_t* h_totals);
int dspx_sk_res_tscr(dspx_sk* sk, uint8_t* d_mask);
/* SpectralKurtosis::mask with the resolutions (:1465-1566): the output's
   ndat_out samples (dspx_sk_blocks_res) are input samples M_last -
   M_last/noverlap_last on, written in blocks of the first resolution's M
   every M/noverlap samples from mask row (M_last - M_last/noverlap_last) /
   (M/noverlap), each block zeroed where its row is flagged. */
int dspx_sk_mask_res(dspx_sk* sk, const uint8_t* d_mask, const float* d_in, uint64_t in_chan_stride,
                     uint64_t in_pol_stride, float* d_out, uint64_t out_chan_stride, uint64_t out_pol_stride,
                     uint64_t ndat);

/* ------------------------------------------------------ fused LoadToFold */

typedef struct dspx_ltf dspx_ltf;

typedef struct {
  uint32_t nbit;          /* 8 or 16: complex TFP [ndat][nchan][npol][2] */
  uint32_t offset_binary;
  float scale;            /* 8-bit unpack scale */
  uint32_t nchan;
  uint32_t npol;          /* 1 or 2 */
  uint32_t reserved0;
  uint64_t ndat;          /* input samples per channel and polarisation */
  dspx_dedisp_desc dedisp;
  uint32_t nfft, nfilt_pos, nfilt_neg;
  uint32_t chirp_mode;    /* DSPX_CHIRP_PARITY or DSPX_CHIRP_TRUTH */
  uint32_t state;         /* DSPX_INTENSITY ... DSPX_STOKES */
  uint32_t ndim;          /* 1, 2 or 4 for Coherence / Stokes */
  uint32_t nbin;
  uint32_t form;          /* DSPX_FORM_* */
} dspx_ltf_desc;

int dspx_ltf_create(dspx_ctx* ctx, const dspx_ltf_desc* desc, dspx_ltf** out);
void dspx_ltf_destroy(dspx_ltf* ltf);
/* Number of FFT parts and output samples of one block (Convolution.C:485-487,
   601-608). */
int dspx_ltf_geometry(const dspx_ltf* ltf, uint64_t* npart, uint64_t* ndat_out);
/* Fold plan of the block (output sample indices). */
int dspx_ltf_set_fold(dspx_ltf* ltf, double phase_per_sample, const dspx_fold_chunk* chunks,
                      uint32_t nchunk);
/* Raw device input -> profiles prof[isub][nchan][npol_out][nbin*ndim_out]
   and hits[isub][nbin], both accumulated. */
int dspx_ltf_execute(dspx_ltf* ltf, const void* d_raw, float* d_prof, uint32_t* d_hits);
/* The SpectralKurtosis-excised chain (LoadToFold1.C:265-377 with
   coherent dedispersion: SpectralKurtosis on the zero-DM copy, its mask on
   the dedispersed samples, SpectralKurtosis.C:1465-1566, then Detection and
   Fold of zeroed data, Fold.C:911-930).  d_mask [nblk][nchan] bytes is the
   dspx_sk_detect mask over the block's output samples in blocks of M (the
   zero-DM input is the raw block from input sample nfilt_pos,
   Convolution.C:792-799: dspx_sk_compute_raw with t0 = nfilt_pos and
   nblk <= ndat_out / M blocks).  Output samples of flagged blocks, and those
   past nblk*M (not in the cleaned series), are zeroed: they add nothing to
   the profiles and no hits.  Hits are per channel, hits[isub][nchan][nbin]
   (hits_nchan = nchan), counting the samples whose first detected value is
   non-zero, accumulated; their sum is Fold's ndat_folded.  M >= 2;
   ndat_out + nfft < 2^31. */
int dspx_ltf_execute_masked(dspx_ltf* ltf, const void* d_raw, const uint8_t* d_mask, uint32_t M, uint64_t nblk,
                            float* d_prof, uint32_t* d_hits);
/* Raw device input -> dedispersed filterbank out[ichan][ipol][npart*step]
   complex (what the Convolution operation outputs). */
int dspx_ltf_filterbank(dspx_ltf* ltf, const void* d_raw, float* d_out);
/* Stream mode (InputBuffering::set_next_start / pre_transformation,
   InputBuffering.C:66-180, with TimeSeries::prepend, TimeSeries.C:514-533):
   consecutive calls each bring d.ndat (>= nfft) new raw samples; the
   unprocessed tail of the previous call (the overlap and any remainder) is
   kept on the device and prepended, so the FFT windows tile the stream as
   one long block would.  Fused form only.  reset starts a new stream. */
int dspx_ltf_stream_reset(dspx_ltf* ltf);
/* The next call's FFT parts and output samples, and the stream index of the
   input sample under its first output sample (nfilt_pos after the first
   unprocessed sample, Convolution.C:611).  Call before each stream call, and
   set that call's fold plan (dspx_ltf_set_fold) after it. */
int dspx_ltf_stream_next(dspx_ltf* ltf, uint64_t* npart, uint64_t* ndat_out, uint64_t* t_first);
/* One stream call: profiles / hits as dspx_ltf_execute, or the call's
   filterbank out[ichan][ipol][ndat_out] as dspx_ltf_filterbank. */
int dspx_ltf_stream_execute(dspx_ltf* ltf, const void* d_raw, float* d_prof, uint32_t* d_hits);
int dspx_ltf_stream_filterbank(dspx_ltf* ltf, const void* d_raw, float* d_out);
/* Per-kernel durations (ms) of the last execute when timing is enabled;
   names[i] points at static strings.  Returns the count via *n. */
int dspx_ltf_set_timing(dspx_ltf* ltf, int enable);
int dspx_ltf_stage_times(dspx_ltf* ltf, float* ms, const char** names, int cap, int* n);

/* ------------------------------------------------------------------- batch */
/* Many beams of uneven shape (channel count, DM, band, FFT length, 8/16-bit
   input, fold bins) through the fused LoadToFold path as one unit of work:
   one plan per beam, one set of large scratch buffers shared by all (sized
   for the largest beam), the whole sequence captured once into a HIP graph
   and replayed.  No reference counterpart: dspsr runs one LoadToFold
   pipeline per beam (LoadToFold1.C:109-258); the FFT-length choice replaces
   the cost model of optimal_fft_length (optimize_fft.c:64-135, used by
   Response::set_optimal_ndat, Response.C:553-588) when autotuned. */
typedef struct dspx_batch dspx_batch;

enum {
  DSPX_BATCH_AUTOTUNE = 1,
  /* bits 8-11: the number of branches (0 or 1: every beam on the context's
     stream; k: beams split over k streams, the longest beams first to the
     least loaded branch, each branch with scratch of its own; executes
     fork from the context's stream and join it again) */
  DSPX_BATCH_BRANCH_SHIFT = 8,
  DSPX_BATCH_BRANCH_MASK = 0xf00
};

/* beams[i] as for dspx_ltf_create (fused form).  A beam with nfft = 0 gets
   its FFT length and overlap from the reference's rule (raised to 4096, the
   fused form's shortest transform, where the rule gives less), or, with
   DSPX_BATCH_AUTOTUNE, the fastest per output sample of the reference's
   length and the next two powers of two (timed on the device here). */
int dspx_batch_create(dspx_ctx* ctx, const dspx_ltf_desc* beams, uint32_t nbeam, uint32_t flags,
                      dspx_batch** out);
void dspx_batch_destroy(dspx_batch* batch);
/* The plan of beam ibeam: its descriptor with the chosen nfft / overlap and
   the form it runs (DSPX_FORM_ONEPASS or DSPX_FORM_TWOLEVEL), FFT parts and
   output samples per block. */
int dspx_batch_beam(const dspx_batch* batch, uint32_t ibeam, dspx_ltf_desc* desc, uint64_t* npart,
                    uint64_t* ndat_out);
/* As dspx_ltf_set_fold for beam ibeam (before capture). */
int dspx_batch_set_fold(dspx_batch* batch, uint32_t ibeam, double phase_per_sample, const dspx_fold_chunk* chunks,
                        uint32_t nchunk);
/* Every beam once, in order: d_raw[i] -> d_prof[i], d_hits[i] (d_hits may be
   NULL), as dspx_ltf_execute. */
int dspx_batch_execute(dspx_batch* batch, const void* const* d_raw, float* const* d_prof, uint32_t* const* d_hits);
/* Captures dspx_batch_execute on these buffers into a graph (the buffers are
   baked in; the context's stream must not be used by other work meanwhile). */
int dspx_batch_capture(dspx_batch* batch, const void* const* d_raw, float* const* d_prof, uint32_t* const* d_hits);
/* Replays the captured graph on the context's stream. */
int dspx_batch_replay(dspx_batch* batch);

/* ------------------------------------------------------- RFI proposal net */
/* No reference counterpart: a learned proposal of RFI-affected blocks ahead
   of the spectral kurtosis detector (which keeps the final decision and
   stays the fallback).  Features per (block, channel) from the SK sums; a 4-layer 1-D
   convolutional classifier across frequency channels (4 -> 32 -> 32 -> 32
   -> 1, 19 taps, ReLU), BF16-rounded weights with FP32 accumulation and an
   FP32 master copy; class-weighted logistic loss; Adam.  Features and
   logits are [nblk][nfeat][nchan] and [nblk][nchan] floats. */
typedef struct dspx_rfinet dspx_rfinet;

/* Number of parameters (weights W[co][ci][k] then biases, layer by layer). */
uint32_t dspx_rfinet_nparam(void);
/* He-uniform initial weights from seed, zero biases. */
int dspx_rfinet_create(dspx_ctx* ctx, uint64_t seed, dspx_rfinet** out);
void dspx_rfinet_destroy(dspx_rfinet* net);
/* Host copies of the FP32 master weights (rounded = 0) or of the BF16-rounded
   weights the passes use (rounded = 1).  Synchronous. */
int dspx_rfinet_get_params(dspx_rfinet* net, float* host, int rounded);
int dspx_rfinet_set_params(dspx_rfinet* net, const float* host);
/* Features [nblk][4][nchan] from SK sums [nblk][nchan][npol][2] and counts:
   log mean power minus its channel mean over the blocks, the mean and the
   largest |SK - 1| of the pols, the change of log power since the previous
   block.  d_scratch: 3*nblk*nchan + nchan floats. */
int dspx_rfi_features(dspx_ctx* ctx, const float* d_sums, const uint32_t* d_counts, uint64_t nblk, uint32_t nchan,
                      uint32_t npol, uint32_t Nd, float* d_scratch, float* d_feat);
int dspx_rfinet_forward(dspx_rfinet* net, const float* d_feat, uint64_t nblk, uint32_t nchan, float* d_logits);
/* Forward, loss (written to d_loss if not NULL) and the gradient of the mean
   class-weighted logistic loss against labels [nblk][nchan] (bytes, 1 =
   flagged) with respect to every parameter (kept on the device). */
int dspx_rfinet_backward(dspx_rfinet* net, const float* d_feat, const uint8_t* d_label, uint64_t nblk,
                         uint32_t nchan, float pos_weight, float* d_loss);
/* Host copy of the last gradient.  Synchronous. */
int dspx_rfinet_grad(dspx_rfinet* net, float* host);
/* One Adam step (step number t >= 1) with the last gradient. */
int dspx_rfinet_adam(dspx_rfinet* net, float lr, float beta1, float beta2, float eps, uint32_t t);
/* mask[i] = logits[i] > threshold */
int dspx_rfinet_propose(dspx_ctx* ctx, const float* d_logits, uint64_t n, float threshold, uint8_t* d_mask);

#ifdef __cplusplus
}
#endif

#endif /* DSPX_H */
