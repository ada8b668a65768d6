// ic.cpp -- seeded initial conditions, draw for draw as the reference's
// generate_particle (src/main.cpp:184-275) with std::default_random_engine
// (minstd_rand0 in libstdc++) and freshly constructed distributions per
// draw (random_datatype builds its distribution object locally).
//
// Order of draws per particle: mass; then
//   cube:   position z, y, x (g++ evaluates the three constructor arguments
//           right to left, src/main.cpp:222); velocity direction z, y, x;
//           speed;
//   sphere: radius fraction (cube root taken), azimuth, cos(polar);
//   shell:  azimuth, cos(polar) -- plus, with the shell_jitter extension
//           (0 = reference behaviour), a radius factor drawn after the mass.
#include <cmath>
#include <random>

#include "../../../include/nbs.h"
#include "../../../include/nbs_host.h"

namespace {

template <class T>
struct Gen {
  std::minstd_rand0 eng;
  bool bad = false;

  T draw(T low, T high, uint32_t dist) {  // src/main.cpp:184-218
    T r = -1;
    switch (dist) {
      case NBS_DIST_LINEAR: {
        std::uniform_real_distribution<T> d(low, high);
        r = d(eng);
        break;
      }
      case NBS_DIST_EXP: {
        const T lambda = (T)(-1.0 * std::log(0.001) / (high - low));
        std::exponential_distribution<T> d(lambda);
        do {
          r = d(eng) + low;
        } while (r > high || r < low);
        break;
      }
      case NBS_DIST_NORMAL: {
        const T centre = (T)((high + low) / 2.0);
        const T dev = (T)((high - low) / 6.0);
        std::normal_distribution<T> d(centre, dev);
        do {
          r = d(eng);
        } while (r > high || r < low);
        break;
      }
      default:
        bad = true;
    }
    return r;
  }

  void triple(T low, T high, T v[3]) {
    v[2] = draw(low, high, NBS_DIST_LINEAR);
    v[1] = draw(low, high, NBS_DIST_LINEAR);
    v[0] = draw(low, high, NBS_DIST_LINEAR);
  }
};

// |v| as src/vector.cpp:149-152 computes it (squares in double).
inline double mag(const double v[3]) {
  return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}
inline float mag(const float v[3]) {
  const double a = v[0], b = v[1], c = v[2];
  return (float)std::sqrt(a * a + b * b + c * c);
}

template <class T>
int generate(const nbs_settings* s, unsigned long seed, size_t n, double* pos,
             double* vel, double* mass) {
  Gen<T> g;
  g.eng.seed(seed);
  T rot[3] = {(T)s->rotation_vector[0], (T)s->rotation_vector[1],
              (T)s->rotation_vector[2]};
  {
    const T m = mag(rot);
    for (int k = 0; k < 3; k++) rot[k] /= m;
    for (int k = 0; k < 3; k++) rot[k] *= (T)s->rotation_magnitude;
  }
  const T scale[3] = {(T)s->scale_x, (T)s->scale_y, (T)s->scale_z};
  for (size_t i = 0; i < n; i++) {
    const T m = g.draw((T)s->min_mass, (T)s->max_mass, s->mass_dist);
    T p[3], v[3];
    if (s->gen_type == NBS_GEN_CUBE) {
      const T sz = (T)s->size;
      g.triple((T)(-1 * (sz / 2)), (T)(sz / 2), p);
      for (int k = 0; k < 3; k++) p[k] *= scale[k];
      g.triple((T)-1, (T)1, v);
      const T vm = mag(v);
      for (int k = 0; k < 3; k++) v[k] /= vm;
      const T speed = g.draw((T)s->min_vel, (T)s->max_vel, s->vel_dist);
      for (int k = 0; k < 3; k++) v[k] *= speed;
    } else if (s->gen_type == NBS_GEN_SPHERE || s->gen_type == NBS_GEN_SHELL) {
      T radius;
      if (s->gen_type == NBS_GEN_SPHERE) {
        radius = g.draw((T)0, (T)1, NBS_DIST_LINEAR);
        radius = std::cbrt(radius) * (T)s->r_sphere;
      } else {
        radius = (T)s->r_sphere;
        if (s->shell_jitter != 0)
          radius = radius * ((T)1 + (T)s->shell_jitter *
                                        g.draw((T)-1, (T)1, NBS_DIST_LINEAR));
      }
      const T az = g.draw((T)0, (T)6.28318530718, NBS_DIST_LINEAR);
      const T polar = std::acos(g.draw((T)-1, (T)1, NBS_DIST_LINEAR));
      p[0] = radius * std::sin(polar) * std::cos(az);
      p[1] = radius * std::sin(polar) * std::sin(az);
      p[2] = radius * std::cos(polar);
      for (int k = 0; k < 3; k++) p[k] *= scale[k];
      // cross(rotation, point), src/vector.cpp:162-165
      v[0] = rot[1] * p[2] - rot[2] * p[1];
      v[1] = rot[2] * p[0] - rot[0] * p[2];
      v[2] = rot[0] * p[1] - rot[1] * p[0];
    } else {
      return NBS_ERR_ARG;
    }
    if (g.bad) return NBS_ERR_ARG;
    for (int k = 0; k < 3; k++) {
      pos[3 * i + k] = p[k];
      vel[3 * i + k] = v[k];
    }
    mass[i] = m;
  }
  return NBS_OK;
}

}  // namespace

extern "C" int nbs_generate(const nbs_settings* s, unsigned long seed,
                            size_t n, int prec32, double* pos, double* vel,
                            double* mass) {
  if (!s || (n && (!pos || !vel || !mass))) return NBS_ERR_ARG;
  return prec32 ? generate<float>(s, seed, n, pos, vel, mass)
                : generate<double>(s, seed, n, pos, vel, mass);
}
