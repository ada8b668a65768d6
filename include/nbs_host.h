/* nbs_host.h -- host-side contract of nbodysim kept by libnbs.so: the
 * settings.cfg reader, the seeded particle generator and the .dat / .txt
 * frame files.  These are the parts of the reference's main.cpp that a
 * drop-in must reproduce exactly (same keys, same draws, same bytes); the
 * step itself is in nbs.h.
 */
#ifndef NBS_HOST_H_
#define NBS_HOST_H_

#include <stddef.h>
#include <stdint.h>

#include "nbs.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { NBS_GEN_CUBE = 0, NBS_GEN_SPHERE = 1, NBS_GEN_SHELL = 2 };
enum { NBS_DIST_LINEAR = 0, NBS_DIST_EXP = 1, NBS_DIST_NORMAL = 2 };
enum { NBS_PROJ_FRONT = 0, NBS_PROJ_SIDE = 1, NBS_PROJ_TOP = 2, NBS_PROJ_ISO = 3 };
enum { NBS_COLOR_BW = 0, NBS_COLOR_HEAT = 1 };

/* Every key of the reference's `settings` struct (src/main.cpp:115-160),
 * with the reference's types widened to double, plus the adaptive-dt keys
 * its README documents (README.md:54-60) and this build's extensions
 * (force_precision, shell_jitter, datatype).  `dump_plaintext`, the
 * README's name for dump_text, is not a key: the reference's parser
 * reports it as an unrecognized variable (src/main.cpp:1300-1302). */
typedef struct nbs_settings {
  int read_existing, use_seed, display_progress, dump_binary, dump_text,
      dump_image, overwrite_data, keep_previous_binary, keep_previous_text,
      verbose, damping, adaptive_brightness, nonlinear_brightness, collide,
      cuda;
  uint32_t num_particles, num_frames, threads, img_w, img_h, projection;
  double size, theta, dt;
  unsigned long seed;
  double min_mass, max_mass, min_vel, max_vel, r_sphere, rotation_magnitude;
  double scale_x, scale_y, scale_z, brightness, scale, collision_range;
  double rotation_vector[3];
  uint32_t gen_type, mass_dist, vel_dist, color;
  double min_node_size;
  /* README adaptive-dt keys (not parsed by the reference's code) */
  int adaptive;
  double min_adaptive_dt, max_pos_change, max_vel_change;
  /* extensions */
  int force_precision;   /* `force_precision f32|f64` (default f64) */
  double shell_jitter;   /* `shell_jitter x`: shell radius R(1 + x U(-1,1)) */
  int datatype_float;    /* `datatype float|double` (default double): the
                            reference's build macro (makefile:3 ships
                            -DFLOAT).  float: particles generated as the
                            FLOAT build draws them and .dat records of
                            40 B, written and resumed from; the step's
                            state stays FP64 either way. */
} nbs_settings;

/* Defaults of src/main.cpp:1312-1357 (plus adaptive defaults of the
 * README: min_adaptive_dt 0.00333, max_pos_change 3, max_vel_change 3). */
void nbs_settings_default(nbs_settings* s);

/* Reads a token-stream config like src/main.cpp:1123-1310.  Messages
 * (unrecognized keys) are appended to msg (may be NULL).  Returns NBS_OK,
 * NBS_ERR_IO if the file is missing, NBS_ERR_ARG for an invalid enum value
 * (the reference asserts). */
int nbs_settings_read(nbs_settings* s, const char* path, char* msg,
                      size_t msg_len);

/* Generates n particles exactly as generate_particle does
 * (src/main.cpp:184-275) from an engine seeded with `seed`
 * (std::default_random_engine, src/main.cpp:1388-1389), in FP64 (the
 * reference's -DDOUBLE build) or, with prec32 != 0, in FP32 (-DFLOAT;
 * values widened exactly).  Outputs are [n][3] / [n]. */
int nbs_generate(const nbs_settings* s, unsigned long seed, size_t n,
                 int prec32, double* pos, double* vel, double* mass);

/* nbs_generate on the device, straight into a context's state (replaces
 * nbs_generate + nbs_set_state + nbs_set_systems): system k holds caller
 * indices [offsets[k], offsets[k+1]) generated from an engine seeded with
 * seeds[k], gen type gen_types[k] (NULL: s->gen_type), every other key
 * from s; one thread per particle reaches its first draw by LCG
 * skip-ahead.  FP64 (-DDOUBLE) generation; mass_dist (and, for cubes,
 * vel_dist) must be linear, since exp / normal draw a variable number of
 * words (NBS_ERR_ARG otherwise).  Masses and cubes are bit-identical to
 * nbs_generate; sphere and shell coordinates go through the device's
 * cbrt, acos, sin, cos and may differ from glibc's in the last bits.
 * Extends src/main.cpp:184-275 to the device; no reference counterpart. */
int nbs_generate_state(nbs_ctx* ctx, const nbs_settings* s, int nsys,
                       const int64_t* offsets, const uint64_t* seeds,
                       const uint32_t* gen_types);

/* "./data/NNNN.dat" / ".txt" names of src/main.cpp:283-291 under dir. */
int nbs_frame_filename(const char* dir, unsigned frame, int binary,
                       char* out, size_t out_len);

/* Binary frame: n raw `particle` records {pos, vel, acc, mass}
 * (src/particle.h:33-36), 80 B (double) or 40 B (prec32) each, no header
 * (src/main.cpp:730-743).  Written to a temporary name and renamed. */
int nbs_write_binary(const char* path, size_t n, const double* pos,
                     const double* vel, const double* acc,
                     const double* mass, int prec32);

/* Text frame: "x, y, z\n" per particle at ostream's default precision 6
 * (src/main.cpp:744-759).  Written to a temporary name and renamed. */
int nbs_write_text(const char* path, size_t n, const double* pos);

/* Record count of a binary frame (file size / record size, floored like
 * src/main.cpp:802), or -1 if it cannot be opened. */
int64_t nbs_binary_count(const char* path, int prec32);

/* Reads up to cap records of a binary frame; returns the count read or -1. */
int64_t nbs_read_binary(const char* path, int prec32, size_t cap, double* pos,
                        double* vel, double* acc, double* mass);

/* Resume point of src/main.cpp:790-811: the highest frame <= num_frames
 * with a binary file in dir; returns that frame + 1, or 0 if none. */
int nbs_resume_frame(const char* dir, unsigned num_frames);

/* The reference's main (src/main.cpp:1359-1504) as a call, used by the
 * nbodysim command: reads settings from cfgfile (NULL: "settings.cfg"),
 * resumes from ./data or generates, then runs the frame loop on the GPU in
 * the current directory, frames written under ./data.  Frame f's state is
 * snapshotted asynchronously (nbs_snapshot_begin) and written by a
 * background thread while later frames are computed.  Returns 0 on
 * success, 1 on any failure (where the reference exits). */
int nbs_run_frames(const char* cfgfile);

/* Writes an 8-bit RGB PNG of w x h pixels from rgb ([h][w][3], top row
 * first, as nbs_render_image returns it): the file write_image saves
 * through FreeImage (src/main.cpp:672).  Written to path.tmp, then renamed. */
int nbs_write_png(const char* path, uint32_t w, uint32_t h, const uint8_t* rgb);

#ifdef __cplusplus
}
#endif

#endif /* NBS_HOST_H_ */
