/*
 * fxcorr_datastream.h — the station data boundary in front of the hot path:
 * a station's frames resident in device memory and, per subintegration, the
 * window the station kernels read with its data time and per-block valid
 * flags (DataStream::calculateControlParams, mpifxcorr/src/datastream.cpp:
 * 376-630, as refined for framed formats by Mk5DataStream, mk5.cpp:146-268),
 * for one buffer that holds the whole stream: the reference's ring of
 * segments collapses to one segment because 288 GB of HBM holds a station's
 * scan.  The result feeds fxg_subint_process / fxg_subint_submit directly.
 */
#ifndef FXCORR_DATASTREAM_H
#define FXCORR_DATASTREAM_H

#include <stdint.h>
#include "fxcorr_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A station's stream as Configuration and mark5access describe it. */
typedef struct {
  int32_t framebytes;            /* frame bytes, header included (after multiplexing)    */
  int32_t payloadbytes;          /* getFramePayloadBytes                                  */
  double  framespersecond;       /* getFramesPerSecond                                    */
  int32_t framegranularity;      /* frames per whole-nanosecond step (mark5access)        */
  int32_t bytespersamplenum;     /* getDBytesPerSampleNum                                  */
  int32_t bytespersampledenom;   /* getDBytesPerSampleDenom                                */
  int32_t numchannels;           /* recorded channels of a band (FFT length / 2)          */
  int32_t blockspersend;
  int32_t pad;
  double  bw_mhz;                /* recorded bandwidth                                    */
  int64_t databytes;             /* getDataBytes: bytes sent per subintegration           */
} fxds_format;

/* Configuration::getDataBytes (configuration.cpp:807-842) for VDIF and
 * Mark5B datastreams of a job: guard_ns is Configuration::getGuardNS. */
int64_t fxds_data_bytes(const fxg_job_desc* job, int datastream, double guard_ns);
/* The format of datastream `datastream` of a job (frames per second from
 * the bandwidth, granularity from the frame rate). */
int fxds_format_of(const fxg_job_desc* job, int datastream, double guard_ns, fxds_format* out);

typedef struct fxds_stream fxds_stream;
int fxds_create(const fxds_format* format, fxds_stream** out);
int fxds_destroy(fxds_stream* s);
/* The resident frames: `bytes` of device memory whose first frame starts at
 * (scan, sec, ns) — bufferinfo scan / scanseconds / scanns. */
int fxds_set_buffer(fxds_stream* s, const uint8_t* device_frames, int64_t bytes, int32_t scan,
                    int32_t sec, int32_t ns);
/* The station input of subintegration (scan, sec, ns): delay_start_us and
 * delay_end_us are the model delays (minus the clock's whole seconds) at the
 * subint start and at start + data span (datastream.cpp:388-399).  out gets
 * the data pointer and length inside the buffer, the data time, and flags
 * (nflags >= ceil(blockspersend / 30) entries); a subintegration the stream
 * cannot serve gets sec = FXG_INVALID_SUBINT (mode.h:232).  The caller
 * fills out->interpolator.  *bufferindex is the window's byte offset. */
int fxds_control(fxds_stream* s, int32_t scan, int32_t sec, int32_t ns, double delay_start_us,
                 double delay_end_us, int32_t model_ok, fxg_station_input* out, int32_t* flags,
                 int32_t nflags, int64_t* bufferindex);

#ifdef __cplusplus
}
#endif
#endif
