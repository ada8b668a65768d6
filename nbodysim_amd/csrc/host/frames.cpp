// frames.cpp -- the reference's frame loop (src/main.cpp:1359-1504) as a
// library call: read settings, seed, resume from the newest ./data/NNNN.dat
// or generate, then per frame run one step on the GPU and write the frame
// files.  Frame f's state is copied to pinned host memory on a second
// stream while frame f+1 is computed, and its files are written by a
// background thread while frame f+2 is computed (the reference's
// dump_threaded); the previous frame's files are removed only after the
// new ones are in place.
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/nbs.h"
#include "../../../include/nbs_host.h"

namespace {

bool exists(const char* p) {
  struct stat st;
  return stat(p, &st) == 0;
}

// One frame's arrays: pos, vel, acc point into the context's pinned
// snapshot buffer ([n][3], caller order), mass is the host copy (masses do
// not change).
struct Frame {
  unsigned frame = 0;
  size_t n = 0;
  const double *pos = nullptr, *vel = nullptr, *acc = nullptr, *mass = nullptr;
  const uint8_t* rgb = nullptr;  // the frame image, when dump_image is set
};

int write_frame(const nbs_settings& s, const Frame& f) {
  char bn[4096], tn[4096], pbn[4096], ptn[4096];
  nbs_frame_filename("./data", f.frame, 1, bn, sizeof(bn));
  nbs_frame_filename("./data", f.frame, 0, tn, sizeof(tn));
  for (const char* p : {bn, tn}) {
    if (!s.overwrite_data && exists(p)) {
      fprintf(stderr, "Overwrite disabled, cannot overwrite %s. Aborting.\n", p);
      return NBS_ERR_IO;
    }
  }
  if (s.dump_binary &&
      nbs_write_binary(bn, f.n, f.pos, f.vel, f.acc, f.mass, s.datatype_float) != NBS_OK) {
    fprintf(stderr, "Cannot write %s\n", bn);
    return NBS_ERR_IO;
  }
  if (s.dump_text && nbs_write_text(tn, f.n, f.pos) != NBS_OK) {
    fprintf(stderr, "Cannot write %s\n", tn);
    return NBS_ERR_IO;
  }
  if (f.rgb) {  // gen_image (src/main.cpp:293-300): ./img/NNNN.png
    char in[64];
    snprintf(in, sizeof(in), "./img/%04u.png", f.frame);
    if (!s.overwrite_data && exists(in)) {
      fprintf(stderr, "Overwrite disabled, cannot overwrite %s. Aborting.\n", in);
      return NBS_ERR_IO;
    }
    if (nbs_write_png(in, s.img_w, s.img_h, f.rgb) != NBS_OK) {
      fprintf(stderr, "Cannot save %s. Exiting.\n", in);
      return NBS_ERR_IO;
    }
  }
  if (f.frame > 0) {
    nbs_frame_filename("./data", f.frame - 1, 1, pbn, sizeof(pbn));
    nbs_frame_filename("./data", f.frame - 1, 0, ptn, sizeof(ptn));
    if (!s.keep_previous_binary && exists(pbn)) unlink(pbn);
    if (!s.keep_previous_text && exists(ptn)) unlink(ptn);
  }
  return NBS_OK;
}

}  // namespace

extern "C" int nbs_run_frames(const char* cfgfile) {
  nbs_settings s;
  nbs_settings_default(&s);
  if (!cfgfile) cfgfile = "settings.cfg";
  printf("Reading config file %s\n", cfgfile);
  char msg[8192];
  int r = nbs_settings_read(&s, cfgfile, msg, sizeof(msg));
  if (msg[0]) fputs(msg, stderr);
  if (r != NBS_OK) return 1;

  const unsigned long seed = s.use_seed ? s.seed : (unsigned long)time(NULL);
  std::vector<double> pos, vel, acc, mass;
  unsigned start_frame = 0;
  if (s.read_existing) {
    start_frame = (unsigned)nbs_resume_frame("./data", s.num_frames);
    if (start_frame == 0) {
      printf("No data to resume from.\n");
    } else {
      char bn[4096];
      nbs_frame_filename("./data", start_frame - 1, 1, bn, sizeof(bn));
      const int64_t n = nbs_binary_count(bn, s.datatype_float);
      if (n < 0) {
        fprintf(stderr, "Cannot read %s\n", bn);
        return 1;
      }
      pos.resize(3 * n);
      vel.resize(3 * n);
      acc.resize(3 * n);
      mass.resize(n);
      if (nbs_read_binary(bn, s.datatype_float, n, pos.data(), vel.data(), acc.data(),
                          mass.data()) != n) {
        fprintf(stderr, "Cannot read %s\n", bn);
        return 1;
      }
      printf("Resuming from %u\n", start_frame);
    }
  }
  if (!s.read_existing || start_frame == 0) {
    printf("Seed: %lu\n", seed);
    printf("Generating particles...\n");
    const size_t n = s.num_particles;
    pos.assign(3 * n, 0);
    vel.assign(3 * n, 0);
    acc.assign(3 * n, 0);
    mass.assign(n, 0);
    if (nbs_generate(&s, seed, n, s.datatype_float, pos.data(), vel.data(), mass.data())) {
      fprintf(stderr, "Invalid generation settings.\n");
      return 1;
    }
  }

  nbs_config cfg;
  nbs_config_default(&cfg);
  cfg.theta = s.theta;
  cfg.dt = s.dt;
  cfg.damping = s.damping;
  cfg.force_precision = s.force_precision;
  cfg.adaptive = s.adaptive;
  cfg.min_adaptive_dt = s.min_adaptive_dt;
  cfg.max_pos_change = s.max_pos_change;
  cfg.max_vel_change = s.max_vel_change;
  // threads > 1 selects gen_root_threaded's tree (src/main.cpp:1423); its
  // minimum-size merge (apply_minimum_size_threaded) is not reproduced, so
  // with min_node_size in effect the serial tree is built
  const bool min_size_on = s.min_node_size > 0 && s.verbose;
  cfg.threaded_tree = s.threads > 1 && !min_size_on;
  if (s.threads > 1 && min_size_on)
    fprintf(stderr, "min_node_size with threads > 1: building the serial tree\n");
  nbs_ctx* ctx = nullptr;
  if (nbs_create(&cfg, &ctx) != NBS_OK) {
    fprintf(stderr, "nbs_create failed\n");
    return 1;
  }
  const size_t n = mass.size();
  if (nbs_set_state(ctx, n, pos.data(), vel.data(), mass.data()) != NBS_OK) {
    fprintf(stderr, "%s\n", nbs_last_error(ctx));
    nbs_destroy(ctx);
    return 1;
  }
  // the reference's step counts no interactions: only the running total
  // (scalar work in the walk) is kept
  nbs_set_counting(ctx, NBS_COUNT_TOTAL);
  // the reference applies min_node_size only when verbose is set too
  // (src/main.cpp:1425)
  if (s.min_node_size > 0 && s.verbose &&
      nbs_set_min_node_size(ctx, s.min_node_size, nullptr) != NBS_OK) {
    fprintf(stderr, "%s\n", nbs_last_error(ctx));
    nbs_destroy(ctx);
    return 1;
  }
  if (s.collide && nbs_set_collisions(ctx, 1, s.collision_range) != NBS_OK) {
    fprintf(stderr, "%s\n", nbs_last_error(ctx));
    nbs_destroy(ctx);
    return 1;
  }
  if (s.dump_binary || s.dump_text) mkdir("./data", 0755);
  if (s.dump_image) mkdir("./img", 0755);
  const nbs_image_params image = {s.img_w, s.img_h, s.projection, s.color,
                                  s.adaptive_brightness, s.nonlinear_brightness, s.scale,
                                  s.brightness};
  std::vector<uint8_t> img_buf[2];
  // With collisions the masses change: each frame's are read back into one
  // of two buffers (the writer of the frame before reads the other).
  std::vector<double> mass_buf[2];

  unsigned frame = start_frame;
  printf("Frame %u/%u\n", frame, s.num_frames);
  const bool dumping = s.dump_binary || s.dump_text || s.dump_image;
  std::thread writer;
  int write_status = NBS_OK;
  int status = 0;
  Frame out;
  bool pending = false;  // a snapshot of frame `pending_frame` is in flight
  unsigned pending_frame = 0;
  // Hands the finished snapshot to the writer thread.  The writer of the
  // frame before reads the other pinned buffer; it is joined first, and
  // must be before the next nbs_snapshot_begin reuses its buffer.
  auto hand_over = [&]() -> bool {
    Frame f;
    f.frame = pending_frame;
    f.mass = s.collide || (s.min_node_size > 0 && s.verbose) ? mass_buf[pending_frame & 1].data()
                                                             : mass.data();
    f.rgb = s.dump_image ? img_buf[pending_frame & 1].data() : nullptr;
    if (nbs_snapshot_end(ctx, &f.pos, &f.vel, &f.acc, &f.n) != NBS_OK) {
      fprintf(stderr, "%s\n", nbs_last_error(ctx));
      return false;
    }
    pending = false;
    if (writer.joinable()) writer.join();
    if (write_status != NBS_OK) return false;
    if (s.verbose) printf("Dumping data...\n");
    out = f;
    writer = std::thread([&s, &out, &write_status] { write_status = write_frame(s, out); });
    return true;
  };
  while (frame < s.num_frames) {
    if (s.verbose) printf("Stepping...\n");
    if (nbs_step(ctx, 1) != NBS_OK) {
      fprintf(stderr, "step failed: %s\n", nbs_last_error(ctx));
      status = 1;
      break;
    }
    if (dumping) {
      if (pending && !hand_over()) {
        status = 1;
        break;
      }
      if (s.dump_image) {
        std::vector<uint8_t>& ib = img_buf[frame & 1];
        ib.resize(3 * (size_t)s.img_w * s.img_h);
        if (nbs_render_image(ctx, &image, ib.data()) != NBS_OK) {
          fprintf(stderr, "%s\n", nbs_last_error(ctx));
          status = 1;
          break;
        }
      }
      if (s.collide || (s.min_node_size > 0 && s.verbose)) {
        int64_t m = 0;
        std::vector<double>& mb = mass_buf[frame & 1];
        if (nbs_particle_count(ctx, &m, nullptr, 0) == NBS_OK) mb.resize((size_t)m);
        if (nbs_get_mass(ctx, mb.data()) != NBS_OK) {
          fprintf(stderr, "%s\n", nbs_last_error(ctx));
          status = 1;
          break;
        }
      }
      if (nbs_snapshot_begin(ctx) != NBS_OK) {
        fprintf(stderr, "%s\n", nbs_last_error(ctx));
        status = 1;
        break;
      }
      pending = true;
      pending_frame = frame;
    }
    frame++;
    printf("Frame %u/%u\n", frame, s.num_frames);
  }
  if (pending && !hand_over()) status = 1;
  if (writer.joinable()) writer.join();
  nbs_destroy(ctx);
  return status == 0 && write_status == NBS_OK ? 0 : 1;
}
