// nbs_oracle.cpp -- CPU restatement of fosted3/nbodysim's per-frame step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is loaded by the
// tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
// by nothing else.  The product (nbodysim_amd, libnbs.so) never links it.
//
// It restates, in the reference's own arithmetic (same operation order, same
// libm calls, no FMA contraction -- build with -O2 -ffp-contract=off), the
// reference's serial CPU path:
//   * IC generation ........ src/main.cpp:184-275 (random_datatype,
//                            random_vector, generate_particle)
//   * root from bbox+2 ..... src/main.cpp:887-927 (gen_root)
//   * one-per-leaf octree .. src/octree.cpp:54-85 (allocate_child),
//                            :134-196 (add_particle)
//   * threaded root ........ src/main.cpp:817-886 (gen_root_threaded),
//                            src/octree.cpp:230-336 (calc_*_threaded,
//                            inside)
//   * mass / COM ........... src/octree.cpp:203-221, :252-278
//   * BFS walk + gravity ... src/main.cpp:306-333, :335-421
//   * kick-drift update .... src/particle.cpp:66-74
//   * vector arithmetic .... src/vector.cpp:99-132
// templated over the reference's `datatype` (double for -DDOUBLE, float for
// -DFLOAT, src/vector.h:4-13).
//
// Parity pins (what checks each function; see tests/):
//   * tree, mass, COM, update: bit-exact against oracle/_ref (the reference's
//     own octree.cpp/particle.cpp/vector.cpp compiled here) and against the
//     committed golden fixtures made from it; the threaded root's tree
//     against the reference's own inside(), calc_mass_threaded() and
//     calc_com_threaded() driven by the _ref harness.
//   * gravity / walk: bit-exact against the _ref harness walk; theta = 0
//     reduces to the textbook direct sum; two-body closed form; lattice
//     mirror symmetry.
//   * IC generation: libstdc++'s minstd_rand0 known answer (C++ standard
//     [rand.predef]: 10000th draw of a default-seeded engine is 1043618065),
//     an exact-integer restatement of generate_canonical, and a g++
//     argument-evaluation-order probe (right to left: z is drawn first).
//     The reference's generate_particle itself lives in main.cpp, which
//     does not compile here (it needs FreeImage.h); its sphere and shell
//     branch is pinned from the raw engine words by draw accounting (8n /
//     6n words in FP64, 4n / 3n in FP32), by the geometry each draw fixes
//     (radius, polar cosine, azimuth), by v = omega x p, and by planted
//     mistakes (draws exchanged, sin/cos exchanged, p x omega) that those
//     checks catch (tests/test_oracle_pins.py).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <thread>
#include <utility>
#include <vector>

namespace {

enum { CUBE = 0, SPHERE = 1, SHELL = 2 };
enum { LINEAR = 0, EXP = 1, NORMAL = 2 };

template <class T>
struct V3 {
  T x, y, z;
};

// src/vector.cpp:144-152.  `pow(v, 2)` folds to v*v under g++ -O2; in the
// FLOAT build pow(float, int) promotes to double, so the squares and their
// sum are formed in double and the root is rounded back to float.
inline double ref_norm3(double x, double y, double z) {
  return std::sqrt(x * x + y * y + z * z);
}
inline float ref_norm3(float x, float y, float z) {
  double dx = x, dy = y, dz = z;
  return (float)std::sqrt(dx * dx + dy * dy + dz * dz);
}

// src/main.cpp:306-333 with src/vector.cpp:154-160 (normalize).  `self`
// handling is done by the caller.  Returns the acceleration contribution.
template <class T>
inline V3<T> ref_gravity(const V3<T>& com, T node_mass, const V3<T>& pos,
                         bool damping) {
  V3<T> a = {com.x - pos.x, com.y - pos.y, com.z - pos.z};
  T mag = ref_norm3(a.x, a.y, a.z);
  T r_sq;
  if (!damping) {
    // pow(datatype, int): glibc pow in double, result rounded to datatype.
    r_sq = (T)std::pow((double)mag, -2.0);
  } else {
    // 1.0 / (pow(m, 2) + exp(-1 * m)): the square in double, exp in the
    // datatype's own overload (expf for float), the quotient in double.
    T e = std::exp(-1 * mag);
    r_sq = (T)(1.0 / ((double)mag * (double)mag + (double)e));
  }
  T m2 = ref_norm3(a.x, a.y, a.z);
  a.x /= m2;
  a.y /= m2;
  a.z /= m2;
  a.x *= node_mass;
  a.y *= node_mass;
  a.z *= node_mass;
  const T G = (T)6.67384e-11;
  a.x *= G;
  a.y *= G;
  a.z *= G;
  a.x *= r_sq;
  a.y *= r_sq;
  a.z *= r_sq;
  return a;
}

// ---------------------------------------------------------------- octree --

template <class T>
struct Tree {
  struct Node {
    V3<T> center, com;
    T side, mass;
    int child[8];
    int p;  // particle index for a leaf, -1 otherwise
    int depth;
    uint64_t khi, klo;  // path digits, level 0 in bits 62..60 of khi
  };
  std::vector<Node> nodes;
  const V3<T>* pos = nullptr;
  const T* mass = nullptr;
  int max_depth_seen = 0;
  bool overflow = false;  // identical positions: the reference recurses forever

  static constexpr int kDepthCap = 1100;

  int make(const V3<T>& c, T side, int depth, uint64_t khi, uint64_t klo) {
    Node n;
    n.center = c;
    n.com = c;  // src/octree.cpp:36
    n.side = side;
    n.mass = 0;
    for (int i = 0; i < 8; i++) n.child[i] = -1;
    n.p = -1;
    n.depth = depth;
    n.khi = khi;
    n.klo = klo;
    nodes.push_back(n);
    if (depth > max_depth_seen) max_depth_seen = depth;
    return (int)nodes.size() - 1;
  }

  static int octant(const V3<T>& p, const V3<T>& c) {  // octree.cpp:153-165
    int i = 0;
    if (p.x > c.x) i += 1;
    if (p.y > c.y) i += 2;
    if (p.z > c.z) i += 4;
    return i;
  }

  // src/octree.cpp:54-85: child centre = centre + (s*half)/2, side = half.
  int alloc_child(int n, int i) {
    T half = nodes[n].side / 2;
    int sx = (i & 1) ? 1 : -1, sy = (i & 2) ? 1 : -1, sz = (i & 4) ? 1 : -1;
    V3<T> c = nodes[n].center;
    c.x += sx * half / 2;
    c.y += sy * half / 2;
    c.z += sz * half / 2;
    int d = nodes[n].depth;
    uint64_t khi = nodes[n].khi, klo = nodes[n].klo;
    if (d < 21)
      khi |= (uint64_t)i << (60 - 3 * d);
    else if (d < 42)
      klo |= (uint64_t)i << (60 - 3 * (d - 21));
    int c_idx = make(c, half, d + 1, khi, klo);
    nodes[n].child[i] = c_idx;
    return c_idx;
  }

  // src/octree.cpp:134-196, written as a loop.  A leaf that receives a
  // second particle pushes its resident down first, then retries.
  void add(int n, int par) {
    for (;;) {
      if (nodes[n].depth > kDepthCap) {
        overflow = true;
        return;
      }
      if (nodes[n].p < 0) {
        bool leaf = true;
        for (int i = 0; i < 8; i++)
          if (nodes[n].child[i] >= 0) {
            leaf = false;
            break;
          }
        if (leaf) {
          nodes[n].p = par;
          return;
        }
        int i = octant(pos[par], nodes[n].center);
        int c = nodes[n].child[i];
        if (c < 0) c = alloc_child(n, i);
        n = c;
      } else {
        int q = nodes[n].p;
        int i = octant(pos[q], nodes[n].center);
        int c = nodes[n].child[i];
        if (c < 0) c = alloc_child(n, i);
        add(c, q);  // lands in an empty child: one level of recursion
        nodes[n].p = -1;
      }
    }
  }

  // src/main.cpp:887-927 (gen_root): bbox midpoint, side = largest extent
  // + 2 with the reference's tie-break falling through to the z extent.
  static void root_of(const V3<T>* p, size_t n, V3<T>& origin, T& size) {
    T min_x = std::numeric_limits<T>::max(), min_y = min_x, min_z = min_x;
    T max_x = std::numeric_limits<T>::lowest(), max_y = max_x, max_z = max_x;
    for (size_t i = 0; i < n; i++) {
      if (p[i].x < min_x) min_x = p[i].x;
      if (p[i].x > max_x) max_x = p[i].x;
      if (p[i].y < min_y) min_y = p[i].y;
      if (p[i].y > max_y) max_y = p[i].y;
      if (p[i].z < min_z) min_z = p[i].z;
      if (p[i].z > max_z) max_z = p[i].z;
    }
    origin.x = (T)((min_x + max_x) / 2.0);
    origin.y = (T)((min_y + max_y) / 2.0);
    origin.z = (T)((min_z + max_z) / 2.0);
    if ((max_x - min_x) > (max_y - min_y) && (max_x - min_x) > (max_z - min_z))
      size = (T)((max_x - min_x) + 2.0);
    else if ((max_y - min_y) > (max_x - min_x) &&
             (max_y - min_y) > (max_z - min_z))
      size = (T)((max_y - min_y) + 2.0);
    else
      size = (T)((max_z - min_z) + 2.0);
  }

  void build(const V3<T>* p, const T* m, size_t n, const V3<T>& origin,
             T size) {
    pos = p;
    mass = m;
    nodes.clear();
    nodes.reserve(2 * n + 16);
    make(origin, size, 0, 0, 0);
    for (size_t i = 0; i < n && !overflow; i++) add(0, (int)i);
  }

  // src/octree.cpp:320-336 (inside): inclusive bounds, the offset formed
  // in the datatype and compared with side / 2.0 and (-1.0 * side) / 2.0
  // in double.
  bool inside(int n, const V3<T>& p) const {
    const Node& nd = nodes[n];
    const double hi = nd.side / 2.0, lo = (-1.0 * nd.side) / 2.0;
    const T dx = p.x - nd.center.x, dy = p.y - nd.center.y, dz = p.z - nd.center.z;
    if ((double)dx > hi || (double)dx < lo) return false;
    if ((double)dy > hi || (double)dy < lo) return false;
    if ((double)dz > hi || (double)dz < lo) return false;
    return true;
  }

  // src/main.cpp:817-886 (gen_root_threaded / gen_root_thread): the root
  // allocates all eight children first, then child i receives, in set
  // order, every particle inside() it -- a particle on the root's centre
  // planes goes into every child whose closed cell holds it, one outside
  // every child into none.
  void build_threaded(const V3<T>* p, const T* m, size_t n, const V3<T>& origin, T size) {
    pos = p;
    mass = m;
    nodes.clear();
    nodes.reserve(2 * n + 16);
    make(origin, size, 0, 0, 0);
    for (int i = 0; i < 8; i++) alloc_child(0, i);
    for (int i = 0; i < 8 && !overflow; i++) {
      const int c = nodes[0].child[i];
      for (size_t j = 0; j < n && !overflow; j++)
        if (inside(c, p[j])) add(c, (int)j);
    }
  }

  // The root's COM by calc_com_threaded (src/octree.cpp:286-313): the
  // children's COM times mass summed in child order, then divided by the
  // root's mass -- with no test for a zero mass.
  void root_com_threaded() {
    V3<T> t = {0, 0, 0};
    for (int i = 0; i < 8; i++) {
      const int c = nodes[0].child[i];
      if (c < 0) continue;
      V3<T> t2 = nodes[c].com;
      const T cm = nodes[c].mass;
      t2.x *= cm;
      t2.y *= cm;
      t2.z *= cm;
      t.x += t2.x;
      t.y += t2.y;
      t.z += t2.z;
    }
    const T M = nodes[0].mass;
    t.x /= M;
    t.y /= M;
    t.z /= M;
    nodes[0].com = t;
  }

  size_t leaves() const {
    size_t k = 0;
    for (const Node& nd : nodes) k += nd.p >= 0;
    return k;
  }

  // src/octree.cpp:203-221 (calc_mass) then :252-278 (calc_com), restated
  // as explicit post-order passes; child order 0..7 fixes the rounding.
  // With the threaded tree the root's mass is the same child-order sum
  // (calc_mass_threaded, :230-250) and its COM root_com_threaded's.
  void mass_com(bool threaded = false) {
    // children always have larger indices than their parent (allocation
    // order), so a reverse sweep is a valid post-order for calc_mass.
    for (size_t k = nodes.size(); k-- > 0;) {
      Node& nd = nodes[k];
      if (nd.p >= 0) {
        nd.mass = mass[nd.p];
      } else {
        nd.mass = 0;
        for (int i = 0; i < 8; i++)
          if (nd.child[i] >= 0) nd.mass += nodes[nd.child[i]].mass;
      }
    }
    if (!threaded) {
      com_rec(0);
      return;
    }
    for (int i = 0; i < 8; i++)
      if (nodes[0].child[i] >= 0) com_rec(nodes[0].child[i]);
    root_com_threaded();
  }
  void com_rec(int k) {
    Node& nd = nodes[k];
    if (nd.p >= 0) {
      nd.com = pos[nd.p];
    } else if (nd.mass == 0) {
      nd.com = V3<T>{0, 0, 0};  // children keep their constructor COM
    } else {
      V3<T> t = {0, 0, 0};
      for (int i = 0; i < 8; i++) {
        int c = nodes[k].child[i];
        if (c < 0) continue;
        com_rec(c);
        V3<T> t2 = nodes[c].com;
        T cm = nodes[c].mass;
        t2.x *= cm;
        t2.y *= cm;
        t2.z *= cm;
        t.x += t2.x;
        t.y += t2.y;
        t.z += t2.z;
      }
      T M = nodes[k].mass;
      t.x /= M;
      t.y /= M;
      t.z /= M;
      nodes[k].com = t;
    }
  }

  // src/main.cpp:335-421 for one particle: BFS with a FIFO, MAC
  // side / distance(com, x) > theta on particle-free nodes, self skipped.
  V3<T> walk(int self, T theta, bool damping, std::vector<int>& q,
             long* interactions, double range = -1.0,
             std::vector<std::pair<int, int>>* pairs = nullptr) const {
    V3<T> acc = {0, 0, 0};
    const V3<T>& x = pos[self];
    q.clear();
    q.push_back(0);
    long cnt = 0;
    for (size_t h = 0; h < q.size(); h++) {
      const Node& nd = nodes[q[h]];
      T dist = ref_norm3(nd.com.x - x.x, nd.com.y - x.y, nd.com.z - x.z);
      if (nd.side / dist > theta && nd.p < 0) {
        for (int i = 0; i < 8; i++)
          if (nd.child[i] >= 0) q.push_back(nd.child[i]);
      } else if (nd.p != self) {
        // src/main.cpp:386-396: a leaf's particle closer than the
        // collision range, distance(node particle, particle)
        if (pairs && nd.p >= 0) {
          const V3<T>& o = pos[nd.p];
          if (ref_norm3(o.x - x.x, o.y - x.y, o.z - x.z) < range)
            pairs->push_back({std::min(self, nd.p), std::max(self, nd.p)});
        }
        V3<T> g = ref_gravity(nd.com, nd.mass, x, damping);
        acc.x += g.x;
        acc.y += g.y;
        acc.z += g.z;
        cnt++;
      }
    }
    if (interactions) *interactions = cnt;
    return acc;
  }
};

template <class T>
void load_state(size_t n, const double* pos, const double* mass,
                std::vector<V3<T>>& p, std::vector<T>& m) {
  p.resize(n);
  m.resize(n);
  for (size_t i = 0; i < n; i++) {
    p[i] = V3<T>{(T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2]};
    m[i] = mass ? (T)mass[i] : (T)0;
  }
}

template <class Fn>
void parallel_for(size_t n, int threads, Fn fn) {
  if (threads <= 1 || n < 2) {
    for (size_t i = 0; i < n; i++) fn(i, 0);
    return;
  }
  std::atomic<size_t> next(0);
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; t++)
    ts.emplace_back([&, t] {
      for (;;) {
        size_t b = next.fetch_add(64);
        if (b >= n) break;
        size_t e = std::min(n, b + 64);
        for (size_t i = b; i < e; i++) fn(i, t);
      }
    });
  for (auto& th : ts) th.join();
}

// --------------------------------------------------- IC generation (host) --
// src/main.cpp:184-275.  std::default_random_engine is minstd_rand0 in
// libstdc++ 11; the distributions are constructed per call exactly like
// random_datatype does, so no state carries over between draws.
template <class T>
T random_datatype(T low, T high, int dist, std::minstd_rand0& g, int* err) {
  T r = -1;
  if (dist == LINEAR) {
    std::uniform_real_distribution<T> d(low, high);
    r = d(g);
  } else if (dist == EXP) {
    T lambda = (T)(-1.0 * std::log(0.001) / (high - low));
    std::exponential_distribution<T> d(lambda);
    do {
      r = d(g) + low;
    } while (r > high || r < low);
  } else if (dist == NORMAL) {
    T center = (T)((high + low) / 2.0);
    T deviation = (T)((high - low) / 6.0);
    std::normal_distribution<T> d(center, deviation);
    do {
      r = d(g);
    } while (r > high || r < low);
  } else {
    *err = 1;
  }
  return r;
}

// g++ evaluates the three constructor arguments right to left, so the z
// component is drawn first (src/main.cpp:222).
template <class T>
V3<T> random_vector(T low, T high, std::minstd_rand0& g, int* err) {
  T z = random_datatype(low, high, LINEAR, g, err);
  T y = random_datatype(low, high, LINEAR, g, err);
  T x = random_datatype(low, high, LINEAR, g, err);
  return V3<T>{x, y, z};
}

template <class T>
void normalize(V3<T>& v) {
  T m = ref_norm3(v.x, v.y, v.z);
  v.x /= m;
  v.y /= m;
  v.z /= m;
}

}  // namespace

extern "C" {

struct orc_gen_settings {
  int gen_type, mass_dist, vel_dist, pad;
  double size, min_mass, max_mass, min_vel, max_vel, r_sphere;
  double rotation_magnitude, rot_x, rot_y, rot_z, scale_x, scale_y, scale_z;
};

// Generates n particles in the reference's order from `seed` (the value
// passed to generator.seed, src/main.cpp:1388-1389).  prec 0 = double,
// 1 = float (values widened exactly into the double outputs).
int orc_generate(int prec, const orc_gen_settings* s, unsigned long seed,
                 size_t n, double* pos, double* vel, double* mass) {
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    std::minstd_rand0 g;
    g.seed(seed);
    int err = 0;
    V3<T> rot = {(T)s->rot_x, (T)s->rot_y, (T)s->rot_z};
    for (size_t i = 0; i < n; i++) {
      T m = random_datatype((T)s->min_mass, (T)s->max_mass, s->mass_dist, g,
                            &err);
      V3<T> p, v;
      if (s->gen_type == CUBE) {
        T sz = (T)s->size;
        p = random_vector((T)(-1 * (sz / 2)), (T)(sz / 2), g, &err);
        p.x *= (T)s->scale_x;
        p.y *= (T)s->scale_y;
        p.z *= (T)s->scale_z;
        v = random_vector((T)-1, (T)1, g, &err);
        normalize(v);
        T sp = random_datatype((T)s->min_vel, (T)s->max_vel, s->vel_dist, g,
                               &err);
        v.x *= sp;
        v.y *= sp;
        v.z *= sp;
      } else if (s->gen_type == SPHERE || s->gen_type == SHELL) {
        T radius;
        if (s->gen_type == SPHERE) {
          radius = random_datatype((T)0, (T)1, LINEAR, g, &err);
          radius = std::cbrt(radius) * (T)s->r_sphere;
        } else {
          radius = (T)s->r_sphere;
        }
        T th = random_datatype((T)0, (T)6.28318530718, LINEAR, g, &err);
        T az = std::acos(random_datatype((T)-1, (T)1, LINEAR, g, &err));
        p = V3<T>{radius * std::sin(az) * std::cos(th),
                  radius * std::sin(az) * std::sin(th),
                  radius * std::cos(az)};
        p.x *= (T)s->scale_x;
        p.y *= (T)s->scale_y;
        p.z *= (T)s->scale_z;
        V3<T> r = rot;
        normalize(r);
        r.x *= (T)s->rotation_magnitude;
        r.y *= (T)s->rotation_magnitude;
        r.z *= (T)s->rotation_magnitude;
        // cross(rotation, point), src/vector.cpp:162-165
        v = V3<T>{r.y * p.z - r.z * p.y, r.z * p.x - r.x * p.z,
                  r.x * p.y - r.y * p.x};
      } else {
        return 2;
      }
      if (err) return 3;
      pos[3 * i] = p.x;
      pos[3 * i + 1] = p.y;
      pos[3 * i + 2] = p.z;
      vel[3 * i] = v.x;
      vel[3 * i + 1] = v.y;
      vel[3 * i + 2] = v.z;
      mass[i] = m;
    }
    return 0;
  };
  return prec == 0 ? run(double()) : run(float());
}

// Draws `count` raw engine outputs after seeding (for known-answer tests).
void orc_minstd_raw(unsigned long seed, int use_default, size_t count,
                    uint64_t* out) {
  std::minstd_rand0 g;
  if (!use_default) g.seed(seed);
  for (size_t i = 0; i < count; i++) out[i] = g();
}

// Uniform doubles/floats exactly as std::uniform_real_distribution<T>(a,b)
// produces them from a freshly seeded engine (generate_canonical path).
void orc_uniform_raw(int prec, unsigned long seed, double a, double b,
                     size_t count, double* out) {
  std::minstd_rand0 g;
  g.seed(seed);
  for (size_t i = 0; i < count; i++) {
    if (prec == 0) {
      std::uniform_real_distribution<double> d(a, b);
      out[i] = d(g);
    } else {
      std::uniform_real_distribution<float> d((float)a, (float)b);
      out[i] = d(g);
    }
  }
}

// Root centre and side from the bounding box (src/main.cpp:887-921).
void orc_root(int prec, size_t n, const double* pos, double* center,
              double* side) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    std::vector<V3<T>> p;
    std::vector<T> m;
    load_state<T>(n, pos, nullptr, p, m);
    V3<T> o;
    T s;
    Tree<T>::root_of(p.data(), n, o, s);
    center[0] = o.x;
    center[1] = o.y;
    center[2] = o.z;
    *side = s;
  };
  if (prec == 0)
    run(double());
  else
    run(float());
}

// Full 42-level child-index paths by the reference's centre chain
// (allocate_child's rounded additions + add_particle's strict '>').
void orc_keys(int prec, size_t n, const double* pos, const double* center,
              double side, uint64_t* khi, uint64_t* klo) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    for (size_t i = 0; i < n; i++) {
      V3<T> p = {(T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2]};
      V3<T> c = {(T)center[0], (T)center[1], (T)center[2]};
      T s = (T)side;
      uint64_t hi = 0, lo = 0;
      for (int d = 0; d < 42; d++) {
        int o = Tree<T>::octant(p, c);
        T half = s / 2;
        int sx = (o & 1) ? 1 : -1, sy = (o & 2) ? 1 : -1,
            sz = (o & 4) ? 1 : -1;
        c.x += sx * half / 2;
        c.y += sy * half / 2;
        c.z += sz * half / 2;
        s = half;
        if (d < 21)
          hi |= (uint64_t)o << (60 - 3 * d);
        else
          lo |= (uint64_t)o << (60 - 3 * (d - 21));
      }
      khi[i] = hi;
      klo[i] = lo;
    }
  };
  if (prec == 0)
    run(double());
  else
    run(float());
}

// ----------------------------------------------------------- tree handle --
struct orc_tree {
  int prec;
  Tree<double> td;
  Tree<float> tf;
  std::vector<V3<double>> pd;
  std::vector<double> md;
  std::vector<V3<float>> pf;
  std::vector<float> mf;
};

// Builds the reference tree.  root_mode 0: bbox + 2 (gen_root); 1: the
// caller's centre/side; 2: bbox + 2 with the tree gen_root_threaded builds
// (threads > 1: eight eager root children, inclusive inside()).  Returns
// null on identical-position overflow.
orc_tree* orc_tree_build(int prec, size_t n, const double* pos,
                         const double* mass, int root_mode,
                         const double* center, double side) {
  orc_tree* t = new orc_tree();
  t->prec = prec;
  auto run = [&](auto& tree, auto& p, auto& m) {
    using T = typename std::remove_reference<decltype(m)>::type::value_type;
    load_state<T>(n, pos, mass, p, m);
    V3<T> o;
    T s;
    if (root_mode != 1) {
      Tree<T>::root_of(p.data(), n, o, s);
    } else {
      o = V3<T>{(T)center[0], (T)center[1], (T)center[2]};
      s = (T)side;
    }
    if (root_mode == 2)
      tree.build_threaded(p.data(), m.data(), n, o, s);
    else
      tree.build(p.data(), m.data(), n, o, s);
    if (!tree.overflow) tree.mass_com(root_mode == 2);
    return !tree.overflow;
  };
  bool ok = prec == 0 ? run(t->td, t->pd, t->md) : run(t->tf, t->pf, t->mf);
  if (!ok) {
    delete t;
    return nullptr;
  }
  return t;
}

void orc_tree_free(orc_tree* t) { delete t; }

size_t orc_tree_size(orc_tree* t) {
  return t->prec == 0 ? t->td.nodes.size() : t->tf.nodes.size();
}

size_t orc_tree_leaves(orc_tree* t) {
  return t->prec == 0 ? t->td.leaves() : t->tf.leaves();
}

int orc_tree_max_depth(orc_tree* t) {
  return t->prec == 0 ? t->td.max_depth_seen : t->tf.max_depth_seen;
}

// Pre-order (children 0..7) dump.  Per node: depth, path (khi, klo), leaf
// particle (or -1), first / count in the DFS leaf order, side, centre, mass,
// COM.  `order` receives the particle indices in DFS leaf order.
void orc_tree_dump(orc_tree* t, int* depth, uint64_t* khi, uint64_t* klo,
                   int* leaf, int64_t* first, int64_t* count, double* side,
                   double* center, double* mass, double* com, int* order) {
  auto run = [&](auto& tree) {
    auto& nodes = tree.nodes;
    size_t out = 0;
    int64_t nleaf = 0;
    std::vector<std::pair<int, int>> st;  // (node, state)
    std::vector<size_t> slot_of(nodes.size());
    st.push_back({0, 0});
    while (!st.empty()) {
      auto [k, state] = st.back();
      st.pop_back();
      const auto& nd = nodes[k];
      if (state == 0) {
        size_t o = out++;
        slot_of[k] = o;
        depth[o] = nd.depth;
        khi[o] = nd.khi;
        klo[o] = nd.klo;
        leaf[o] = nd.p;
        first[o] = nleaf;
        side[o] = nd.side;
        center[3 * o] = nd.center.x;
        center[3 * o + 1] = nd.center.y;
        center[3 * o + 2] = nd.center.z;
        mass[o] = nd.mass;
        com[3 * o] = nd.com.x;
        com[3 * o + 1] = nd.com.y;
        com[3 * o + 2] = nd.com.z;
        if (nd.p >= 0) {
          order[nleaf++] = nd.p;
          count[o] = 1;
        } else {
          st.push_back({k, 1});
          for (int i = 7; i >= 0; i--)
            if (nd.child[i] >= 0) st.push_back({nd.child[i], 0});
        }
      } else {
        size_t o = slot_of[k];
        count[o] = nleaf - first[o];
      }
    }
  };
  if (t->prec == 0)
    run(t->td);
  else
    run(t->tf);
}

// Accelerations of every particle by the reference walk (BFS order of
// accepted nodes), `threads` workers.  interactions may be null.
void orc_tree_forces(orc_tree* t, double theta, int damping, int threads,
                     double* acc, int64_t* interactions) {
  auto run = [&](auto& tree) {
    using T = decltype(tree.nodes[0].side);
    size_t n = t->prec == 0 ? t->pd.size() : t->pf.size();
    std::vector<std::vector<int>> qs(std::max(threads, 1));
    parallel_for(n, threads, [&](size_t i, int w) {
      long c = 0;
      V3<T> a = tree.walk((int)i, (T)theta, damping != 0, qs[w], &c);
      acc[3 * i] = a.x;
      acc[3 * i + 1] = a.y;
      acc[3 * i + 2] = a.z;
      if (interactions) interactions[i] = c;
    });
  };
  if (t->prec == 0)
    run(t->td);
  else
    run(t->tf);
}

// Accelerations of a subset of particles (for the bounded CPU sample).
void orc_tree_forces_subset(orc_tree* t, double theta, int damping,
                            int threads, size_t m, const int64_t* idx,
                            double* acc, int64_t* interactions) {
  auto run = [&](auto& tree) {
    using T = decltype(tree.nodes[0].side);
    std::vector<std::vector<int>> qs(std::max(threads, 1));
    parallel_for(m, threads, [&](size_t k, int w) {
      long c = 0;
      V3<T> a = tree.walk((int)idx[k], (T)theta, damping != 0, qs[w], &c);
      acc[3 * k] = a.x;
      acc[3 * k + 1] = a.y;
      acc[3 * k + 2] = a.z;
      if (interactions) interactions[k] = c;
    });
  };
  if (t->prec == 0)
    run(t->td);
  else
    run(t->tf);
}

// One or more frames of the serial reference loop (src/main.cpp:1420-1492
// without collisions, min-node-size or dumps): gen_root, calc_mass,
// calc_com, forces, update.  Arrays are updated in place; acc holds the
// last frame's accelerations, as the reference's particle records do.
int orc_run(int prec, size_t n, double* pos, double* vel, double* acc,
            const double* mass, double theta, double dt, int damping,
            int frames, int threads, int tree_threaded) {
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    std::vector<V3<T>> p, v, a(n);
    std::vector<T> m, dummy;
    load_state<T>(n, pos, mass, p, m);
    load_state<T>(n, vel, nullptr, v, dummy);
    std::vector<std::vector<int>> qs(std::max(threads, 1));
    for (int f = 0; f < frames; f++) {
      Tree<T> tree;
      V3<T> o;
      T s;
      Tree<T>::root_of(p.data(), n, o, s);
      if (tree_threaded)
        tree.build_threaded(p.data(), m.data(), n, o, s);
      else
        tree.build(p.data(), m.data(), n, o, s);
      if (tree.overflow) return 1;
      tree.mass_com(tree_threaded != 0);
      parallel_for(n, threads, [&](size_t i, int w) {
        a[i] = tree.walk((int)i, (T)theta, damping != 0, qs[w], nullptr);
      });
      // src/particle.cpp:66-74
      T tdt = (T)dt;
      for (size_t i = 0; i < n; i++) {
        V3<T> t1 = a[i];
        t1.x *= tdt;
        t1.y *= tdt;
        t1.z *= tdt;
        v[i].x += t1.x;
        v[i].y += t1.y;
        v[i].z += t1.z;
        V3<T> t2 = v[i];
        t2.x *= tdt;
        t2.y *= tdt;
        t2.z *= tdt;
        p[i].x += t2.x;
        p[i].y += t2.y;
        p[i].z += t2.z;
      }
    }
    for (size_t i = 0; i < n; i++) {
      pos[3 * i] = p[i].x;
      pos[3 * i + 1] = p[i].y;
      pos[3 * i + 2] = p[i].z;
      vel[3 * i] = v[i].x;
      vel[3 * i + 1] = v[i].y;
      vel[3 * i + 2] = v[i].z;
      if (acc) {
        acc[3 * i] = a[i].x;
        acc[3 * i + 1] = a[i].y;
        acc[3 * i + 2] = a[i].z;
      }
    }
    return 0;
  };
  return prec == 0 ? run(double()) : run(float());
}

// Frames with collision merging (src/main.cpp:1420-1492 with
// config.collide, FP64): gen_root, calc_mass, calc_com, the walk recording
// pairs closer than `range` (src/main.cpp:386-396), the merge of every
// connected group of pairs into one particle, then update_all.  Each merge
// is collide() (src/main.cpp:423-430): mass a.m + b.m, position and
// velocity weighted_average(a, b, a.m, b.m) (src/vector.cpp:129-132),
// acceleration zero.  The reference merges the pairs in the order of its
// hash set, which is not reproducible; here each group folds collide over
// its members in increasing index (for a group of two the order does not
// matter: both products and the sum commute).  The merged particle takes
// the place of its lowest-index member and the absorbed ones are removed,
// the others keeping their order.  Arrays are compacted in place; returns
// the new count (-1 on a tree overflow).  pairs_out / removed_out (may be
// null) receive the detections and absorbed particles of all frames.
//
// min_size > 0 adds apply_minimum_size (src/main.cpp:1001-1057) before the
// mass and COM: every largest cell of side < min_size becomes one particle,
// collide() folded along the tree (children 0..7 at each level, a leaf's
// result its particle), in its lowest-index member's place; the tree is
// then built again for the merged particles on the same root.  (The
// reference keeps the collapsed tree instead.  It differs only where a
// cell is left with a single particle below a chain of one-child cells,
// whose centre of mass (x m) / m can differ from x in the last bit.)
int64_t orc_run_collide(size_t n, double* pos, double* vel, double* acc, double* mass,
                        double theta, double dt, int damping, double range, int frames,
                        int threads, int64_t* pairs_out, int64_t* removed_out,
                        double min_size) {
  std::vector<V3<double>> p, v, a(n);
  std::vector<double> m, dummy;
  load_state<double>(n, pos, mass, p, m);
  load_state<double>(n, vel, nullptr, v, dummy);
  std::vector<std::vector<int>> qs(std::max(threads, 1));
  int64_t npairs = 0, nremoved = 0;
  for (int f = 0; f < frames && n > 0; f++) {
    Tree<double> tree;
    V3<double> o;
    double side;
    Tree<double>::root_of(p.data(), n, o, side);
    tree.build(p.data(), m.data(), n, o, side);
    if (tree.overflow) return -1;
    if (min_size > 0) {
      struct Part {
        double m;
        V3<double> x, v;
      };
      auto merge = [](Part& acc, const Part& b) {  // collide(), src/main.cpp:423-430
        const double am = acc.m, bm = b.m;
        acc.x = V3<double>{(acc.x.x * am + b.x.x * bm) / (am + bm),
                           (acc.x.y * am + b.x.y * bm) / (am + bm),
                           (acc.x.z * am + b.x.z * bm) / (am + bm)};
        acc.v = V3<double>{(acc.v.x * am + b.v.x * bm) / (am + bm),
                           (acc.v.y * am + b.v.y * bm) / (am + bm),
                           (acc.v.z * am + b.v.z * bm) / (am + bm)};
        acc.m = am + bm;
      };
      std::vector<char> alive(n, 1);
      std::vector<int> members;
      // the fold over a subtree; false when it holds no particle
      std::function<bool(int, Part&)> fold = [&](int k, Part& r) {
        const auto& nd = tree.nodes[k];
        if (nd.p >= 0) {
          r = Part{m[nd.p], p[nd.p], v[nd.p]};
          members.push_back(nd.p);
          return true;
        }
        bool has = false;
        for (int i = 0; i < 8; i++) {
          Part q;
          if (nd.child[i] >= 0 && fold(nd.child[i], q)) {
            if (has) merge(r, q); else r = q;
            has = true;
          }
        }
        return has;
      };
      std::function<void(int)> visit = [&](int k) {
        const auto& nd = tree.nodes[k];
        if (nd.p >= 0) return;
        if (nd.side < min_size) {  // the largest cell below min_size on this path
          members.clear();
          Part r;
          if (!fold(k, r)) return;
          const int keep = *std::min_element(members.begin(), members.end());
          for (int q : members) alive[q] = 0;
          alive[keep] = 1;
          m[keep] = r.m;
          p[keep] = r.x;
          v[keep] = r.v;
          return;
        }
        for (int i = 0; i < 8; i++)
          if (nd.child[i] >= 0) visit(nd.child[i]);
      };
      visit(0);
      size_t k = 0;
      for (size_t i = 0; i < n; i++) {
        if (!alive[i]) continue;
        p[k] = p[i];
        v[k] = v[i];
        m[k] = m[i];
        k++;
      }
      nremoved += (int64_t)(n - k);
      n = k;
      p.resize(n);
      v.resize(n);
      a.resize(n);
      m.resize(n);
      tree = Tree<double>();
      tree.build(p.data(), m.data(), n, o, side);  // the same root
      if (tree.overflow) return -1;
    }
    tree.mass_com();
    std::vector<std::vector<std::pair<int, int>>> tp(std::max(threads, 1));
    parallel_for(n, threads, [&](size_t i, int w) {
      a[i] = tree.walk((int)i, theta, damping != 0, qs[w], nullptr, range, &tp[w]);
    });
    std::vector<std::pair<int, int>> pairs;
    for (auto& t : tp) pairs.insert(pairs.end(), t.begin(), t.end());
    npairs += (int64_t)pairs.size();
    // groups: union-find with the smaller index as root
    std::vector<int> parent(n);
    for (size_t i = 0; i < n; i++) parent[i] = (int)i;
    auto find = [&](int x) {
      while (parent[x] != x) x = parent[x] = parent[parent[x]];
      return x;
    };
    for (auto& q : pairs) {
      int r1 = find(q.first), r2 = find(q.second);
      if (r1 == r2) continue;
      if (r1 > r2) std::swap(r1, r2);
      parent[r2] = r1;
    }
    std::vector<char> member(n, 0), alive(n, 1);
    for (auto& q : pairs) member[q.first] = member[q.second] = 1;
    std::vector<int> head(n, -1);  // group root -> merged particle's index
    for (size_t i = 0; i < n; i++) {
      if (!member[i]) continue;
      const int r = find((int)i);
      if (head[r] < 0) {
        head[r] = (int)i;  // lowest index of the group: the fold starts here
        continue;
      }
      // collide(acc, particle i), src/main.cpp:423-430
      const int h = head[r];
      const double am = m[h], bm = m[i];
      V3<double>& hp = p[h];
      V3<double>& hv = v[h];
      hp = V3<double>{(hp.x * am + p[i].x * bm) / (am + bm), (hp.y * am + p[i].y * bm) / (am + bm),
                      (hp.z * am + p[i].z * bm) / (am + bm)};
      hv = V3<double>{(hv.x * am + v[i].x * bm) / (am + bm), (hv.y * am + v[i].y * bm) / (am + bm),
                      (hv.z * am + v[i].z * bm) / (am + bm)};
      m[h] = am + bm;
      a[h] = V3<double>{0, 0, 0};
      alive[i] = 0;
    }
    // update_all, src/particle.cpp:66-74
    for (size_t i = 0; i < n; i++) {
      V3<double> t1 = a[i];
      t1.x *= dt;
      t1.y *= dt;
      t1.z *= dt;
      v[i].x += t1.x;
      v[i].y += t1.y;
      v[i].z += t1.z;
      V3<double> t2 = v[i];
      t2.x *= dt;
      t2.y *= dt;
      t2.z *= dt;
      p[i].x += t2.x;
      p[i].y += t2.y;
      p[i].z += t2.z;
    }
    size_t k = 0;
    for (size_t i = 0; i < n; i++) {
      if (!alive[i]) continue;
      p[k] = p[i];
      v[k] = v[i];
      a[k] = a[i];
      m[k] = m[i];
      k++;
    }
    nremoved += (int64_t)(n - k);
    n = k;
    p.resize(n);
    v.resize(n);
    a.resize(n);
    m.resize(n);
  }
  for (size_t i = 0; i < n; i++) {
    pos[3 * i] = p[i].x;
    pos[3 * i + 1] = p[i].y;
    pos[3 * i + 2] = p[i].z;
    vel[3 * i] = v[i].x;
    vel[3 * i + 1] = v[i].y;
    vel[3 * i + 2] = v[i].z;
    if (acc) {
      acc[3 * i] = a[i].x;
      acc[3 * i + 1] = a[i].y;
      acc[3 * i + 2] = a[i].z;
    }
    mass[i] = m[i];
  }
  if (pairs_out) *pairs_out = npairs;
  if (removed_out) *removed_out = nremoved;
  return (int64_t)n;
}

// write_image (src/main.cpp:556-678) for FP64 positions, without
// FreeImage: the same projection, bounds, per-particle accumulation
// temp[x][y] += brightness with the running maximum, nonlinear root and
// colour mapping.  rgb: [h][w][3], top row first (FreeImage's scanline 0
// is the bottom one, so pixel (x, y) lands in row h - 1 - y).
void orc_render_image(size_t n, const double* pos, unsigned img_w, unsigned img_h,
                      unsigned projection, unsigned color, int adaptive, int nonlinear,
                      double scale, double brightness, unsigned char* rgb) {
  auto clamp = [](double a, double x, double b) {  // src/main.cpp:177-182
    if (x < a) return a;
    if (x > b) return b;
    return x;
  };
  std::vector<std::vector<double>> temp(img_w, std::vector<double>(img_h, 0.0));
  double max = 0;
  if (projection == 3) scale *= 2.0;
  for (size_t i = 0; i < n; i++) {
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    double x, y;
    if (projection == 0) {
      x = px + (img_w / 2);
      y = pz + (img_h / 2);
    } else if (projection == 1) {
      x = py + (img_w / 2);
      y = pz + (img_h / 2);
    } else if (projection == 2) {
      x = px + (img_w / 2);
      y = py + (img_h / 2);
    } else {
      x = ((std::sqrt(3) / 2.0) * (px - py) + img_w) / 2.0;
      y = ((-0.5) * (px + py + 2 * pz) + img_h) / 2.0;
    }
    x += (x - (img_w / 2)) * (scale - 1);
    y += (y - (img_h / 2)) * (scale - 1);
    if (x < 0 || y < 0 || x > img_w - 1 || y > img_w - 1) continue;
    if (x != x || y != y) continue;  // (int) of NaN: undefined in the reference
    x = clamp(0, x, img_w - 1);
    y = clamp(0, y, img_h - 1);
    if (projection == 3) y = (img_h - 1) - y;
    temp[(int)x][(int)y] += brightness;
    if (temp[(int)x][(int)y] > max) max = temp[(int)x][(int)y];
  }
  if (nonlinear) {
    max = std::sqrt(max);
    for (unsigned i = 0; i < img_w; i++)
      for (unsigned j = 0; j < img_h; j++) temp[i][j] = std::sqrt(temp[i][j]);
  }
  for (unsigned xi = 0; xi < img_w; xi++) {
    for (unsigned yi = 0; yi < img_h; yi++) {
      unsigned char r = 0, g = 0, b = 0;
      double v = color == 0 ? (adaptive ? clamp(0, temp[xi][yi] * 255.0 / max, 255)
                                        : clamp(0, temp[xi][yi], 255))
                            : (adaptive ? clamp(0, temp[xi][yi] * 511.0 / max, 511)
                                        : clamp(0, temp[xi][yi], 511));
      // (int) NaN is INT_MIN on x86; the reference's byte stores and
      // heat-map clamps make that black
      if (v != v) {
      } else if (color == 0) {
        r = g = b = (unsigned char)(int)v;
      } else {
        int iv = (int)v;
        if (iv < 256) {
          b = (unsigned char)clamp(0, 255 - iv, 255);
          g = (unsigned char)clamp(0, iv, 255);
        } else {
          iv -= 256;
          g = (unsigned char)clamp(0, 255 - iv, 255);
          r = (unsigned char)clamp(0, iv, 255);
        }
      }
      unsigned char* o = rgb + ((size_t)(img_h - 1 - yi) * img_w + xi) * 3;
      o[0] = r;
      o[1] = g;
      o[2] = b;
    }
  }
}

// FP64 direct sum (the accuracy reference), long-double accumulation,
// for the particles listed in idx.  Skips j == i.  damping follows the
// reference's damped law.
void orc_direct(size_t n, const double* pos, const double* mass,
                int damping, size_t m, const int64_t* idx, int threads,
                double* acc) {
  parallel_for(m, threads, [&](size_t k, int) {
    int64_t i = idx[k];
    long double ax = 0, ay = 0, az = 0;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    for (size_t j = 0; j < n; j++) {
      if ((int64_t)j == i) continue;
      long double dx = (long double)pos[3 * j] - xi;
      long double dy = (long double)pos[3 * j + 1] - yi;
      long double dz = (long double)pos[3 * j + 2] - zi;
      long double d2 = dx * dx + dy * dy + dz * dz;
      long double d = sqrtl(d2);
      long double s;
      if (!damping)
        s = 6.67384e-11L * mass[j] / (d2 * d);
      else
        s = 6.67384e-11L * mass[j] / ((d2 + expl(-d)) * d);
      ax += s * dx;
      ay += s * dy;
      az += s * dz;
    }
    acc[3 * k] = (double)ax;
    acc[3 * k + 1] = (double)ay;
    acc[3 * k + 2] = (double)az;
  });
}

}  // extern "C"
