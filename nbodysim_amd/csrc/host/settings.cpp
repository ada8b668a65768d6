// settings.cpp -- settings.cfg reader with the reference's semantics
// (src/main.cpp:1123-1310): whitespace-separated tokens; a boolean is true
// only for the literal "true"; enums outside their set are an error (the
// reference asserts); `rotation_vector x y z` takes three numbers; stray
// "true"/"false" tokens are skipped; anything else is reported as
// "Unrecognized variable".  Numbers are read like istream >> does: the
// longest numeric prefix of the token, the rest stays in the stream (so
// `seed 0xDEADBEEF` reads 0 and then reports "xDEADBEEF").
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <string>

#include "../../../include/nbs.h"
#include "../../../include/nbs_host.h"

namespace {

struct Tokens {
  std::deque<std::string> q;
  bool next(std::string& out) {
    if (q.empty()) return false;
    out = q.front();
    q.pop_front();
    return true;
  }
  // Numeric extraction: consumes the numeric prefix, pushes back the rest.
  template <class Parse>
  bool number(Parse parse) {
    std::string t;
    if (!next(t)) return false;
    size_t used = parse(t);
    if (used == 0) {
      q.push_front(t);  // failed extraction leaves the token (stream fails)
      return false;
    }
    if (used < t.size()) q.push_front(t.substr(used));
    return true;
  }
};

// Decimal floating-point prefix as accepted by num_get (no inf/nan/hex).
size_t float_prefix(const std::string& t) {
  size_t i = 0, n = t.size();
  if (i < n && (t[i] == '+' || t[i] == '-')) i++;
  size_t digits = 0;
  while (i < n && isdigit((unsigned char)t[i])) i++, digits++;
  if (i < n && t[i] == '.') {
    i++;
    while (i < n && isdigit((unsigned char)t[i])) i++, digits++;
  }
  if (digits == 0) return 0;
  if (i < n && (t[i] == 'e' || t[i] == 'E')) {
    size_t j = i + 1;
    if (j < n && (t[j] == '+' || t[j] == '-')) j++;
    size_t ed = 0;
    while (j < n && isdigit((unsigned char)t[j])) j++, ed++;
    if (ed) i = j;
  }
  return i;
}

size_t uint_prefix(const std::string& t) {
  size_t i = 0, n = t.size();
  if (i < n && t[i] == '+') i++;
  size_t d = 0;
  while (i < n && isdigit((unsigned char)t[i])) i++, d++;
  return d ? i : 0;
}

bool read_double(Tokens& tk, double& v) {
  return tk.number([&](const std::string& t) {
    size_t u = float_prefix(t);
    if (u) v = strtod(t.substr(0, u).c_str(), nullptr);
    return u;
  });
}

template <class U>
bool read_unsigned(Tokens& tk, U& v) {
  return tk.number([&](const std::string& t) {
    size_t u = uint_prefix(t);
    if (u) v = (U)strtoull(t.substr(0, u).c_str(), nullptr, 10);
    return u;
  });
}

void note(char* msg, size_t len, const std::string& s) {
  if (!msg || len == 0) return;
  size_t cur = strnlen(msg, len);
  if (cur + 1 >= len) return;
  snprintf(msg + cur, len - cur, "%s\n", s.c_str());
}

}  // namespace

extern "C" {

void nbs_settings_default(nbs_settings* s) {
  std::memset(s, 0, sizeof(*s));
  s->read_existing = 0;
  s->use_seed = 0;
  s->display_progress = 1;
  s->dump_binary = 1;
  s->dump_text = 1;
  s->dump_image = 1;
  s->overwrite_data = 0;
  s->keep_previous_binary = 0;
  s->keep_previous_text = 1;
  s->verbose = 0;
  s->num_particles = 5000;
  s->num_frames = 300;
  s->size = 512;
  s->theta = 0.5;
  s->dt = 0.0333;
  s->seed = 0xDEADBEEFul;
  s->min_mass = 5e10;
  s->max_mass = 5e11;
  s->min_vel = 0;
  s->max_vel = 0;
  s->r_sphere = 100;
  s->rotation_magnitude = 0.1;
  s->rotation_vector[0] = 0;
  s->rotation_vector[1] = 0;
  s->rotation_vector[2] = 1;
  s->gen_type = NBS_GEN_CUBE;
  s->mass_dist = NBS_DIST_LINEAR;
  s->vel_dist = NBS_DIST_LINEAR;
  s->scale_x = 1;
  s->scale_y = 1;
  s->scale_z = 1;
  s->damping = 0;
  s->threads = 1;
  s->brightness = 255;
  s->projection = NBS_PROJ_ISO;
  s->img_w = 1920;
  s->img_h = 1080;
  s->scale = 1;
  s->color = NBS_COLOR_BW;
  s->adaptive_brightness = 0;
  s->nonlinear_brightness = 0;
  s->collide = 0;
  s->collision_range = 0.01;
  s->min_node_size = 0.0;
  s->cuda = 0;
  s->adaptive = 0;
  s->min_adaptive_dt = 0.00333;
  s->max_pos_change = 3;
  s->max_vel_change = 3;
  s->force_precision = NBS_FORCE_F64;
  s->shell_jitter = 0;
  s->datatype_float = 0;
}

int nbs_settings_read(nbs_settings* s, const char* path, char* msg,
                      size_t msg_len) {
  if (msg && msg_len) msg[0] = 0;
  std::ifstream f(path);
  if (!f.good()) {
    note(msg, msg_len, std::string(path) + " doesn't exist or can't be read.");
    return NBS_ERR_IO;
  }
  Tokens tk;
  std::string w;
  while (f >> w) tk.q.push_back(w);

  struct BoolKey {
    const char* name;
    int* dst;
  } bools[] = {{"read_existing", &s->read_existing},
               {"use_seed", &s->use_seed},
               {"display_progress", &s->display_progress},
               {"dump_binary", &s->dump_binary},
               {"dump_text", &s->dump_text},
               {"dump_image", &s->dump_image},
               {"overwrite_data", &s->overwrite_data},
               {"keep_previous_binary", &s->keep_previous_binary},
               {"keep_previous_text", &s->keep_previous_text},
               {"verbose", &s->verbose},
               {"damping", &s->damping},
               {"adaptive_brightness", &s->adaptive_brightness},
               {"nonlinear_brightness", &s->nonlinear_brightness},
               {"collide", &s->collide},
               {"cuda", &s->cuda},
               {"adaptive", &s->adaptive}};
  struct DblKey {
    const char* name;
    double* dst;
  } dbls[] = {{"scale_x", &s->scale_x},
              {"scale_y", &s->scale_y},
              {"scale_z", &s->scale_z},
              {"r_sphere", &s->r_sphere},
              {"rotation_magnitude", &s->rotation_magnitude},
              {"size", &s->size},
              {"theta", &s->theta},
              {"dt", &s->dt},
              {"min_mass", &s->min_mass},
              {"max_mass", &s->max_mass},
              {"min_vel", &s->min_vel},
              {"max_vel", &s->max_vel},
              {"brightness", &s->brightness},
              {"scale", &s->scale},
              {"collision_range", &s->collision_range},
              {"min_node_size", &s->min_node_size},
              {"min_adaptive_dt", &s->min_adaptive_dt},
              {"max_pos_change", &s->max_pos_change},
              {"max_vel_change", &s->max_vel_change},
              {"shell_jitter", &s->shell_jitter}};
  struct UKey {
    const char* name;
    uint32_t* dst;
  } uints[] = {{"threads", &s->threads},
               {"num_particles", &s->num_particles},
               {"num_frames", &s->num_frames},
               {"img_w", &s->img_w},
               {"img_h", &s->img_h}};
  struct EnumKey {
    const char* name;
    uint32_t* dst;
    const char* values[4];
  } enums[] = {{"gen_type", &s->gen_type, {"cube", "sphere", "shell", nullptr}},
               {"mass_dist", &s->mass_dist, {"linear", "exp", "normal", nullptr}},
               {"vel_dist", &s->vel_dist, {"linear", "exp", "normal", nullptr}},
               {"projection", &s->projection, {"front", "side", "top", "iso"}},
               {"color", &s->color, {"bw", "heat", nullptr, nullptr}}};

  std::string var;
  while (tk.next(var)) {
    bool done = false;
    for (auto& b : bools)
      if (var == b.name) {
        std::string v;
        *b.dst = tk.next(v) && v == "true";
        done = true;
        break;
      }
    if (done) continue;
    for (auto& d : dbls)
      if (var == d.name) {
        read_double(tk, *d.dst);
        done = true;
        break;
      }
    if (done) continue;
    for (auto& u : uints)
      if (var == u.name) {
        read_unsigned(tk, *u.dst);
        done = true;
        break;
      }
    if (done) continue;
    for (auto& e : enums)
      if (var == e.name) {
        std::string v;
        tk.next(v);
        int found = -1;
        for (int k = 0; k < 4; k++)
          if (e.values[k] && v == e.values[k]) found = k;
        if (found < 0) {
          note(msg, msg_len, std::string("invalid value '") + v + "' for " + var);
          return NBS_ERR_ARG;
        }
        *e.dst = (uint32_t)found;
        done = true;
        break;
      }
    if (done) continue;
    if (var == "rotation_vector") {
      read_double(tk, s->rotation_vector[0]);
      read_double(tk, s->rotation_vector[1]);
      read_double(tk, s->rotation_vector[2]);
    } else if (var == "seed") {
      read_unsigned(tk, s->seed);
    } else if (var == "force_precision") {
      std::string v;
      tk.next(v);
      if (v == "f32" || v == "float")
        s->force_precision = NBS_FORCE_F32;
      else if (v == "f64" || v == "double")
        s->force_precision = NBS_FORCE_F64;
      else {
        note(msg, msg_len, "invalid value '" + v + "' for force_precision");
        return NBS_ERR_ARG;
      }
    } else if (var == "datatype") {
      // the reference's build macro (-DFLOAT / -DDOUBLE, makefile:3) as a key
      std::string v;
      tk.next(v);
      if (v == "float")
        s->datatype_float = 1;
      else if (v == "double")
        s->datatype_float = 0;
      else {
        note(msg, msg_len, "invalid value '" + v + "' for datatype");
        return NBS_ERR_ARG;
      }
    } else if (var == "true" || var == "false") {
      continue;
    } else {
      note(msg, msg_len, "Unrecognized variable " + var);
    }
  }
  return NBS_OK;
}

}  // extern "C"
