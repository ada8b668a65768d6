// ref_harness.cpp -- driver for the reference's own compiled sources.
//
// TEST INFRASTRUCTURE ONLY (oracle).  oracle/build_ref.sh compiles this file
// together with the reference's src/octree.cpp, src/particle.cpp,
// src/vector.cpp and src/thread_functions.cpp, taken where they lie under
// the reference tree, into oracle/_ref/libnbsref_{f64,f32}.so.  Those four
// files build with g++ alone; src/main.cpp does not (it needs FreeImage.h),
// so the few main.cpp functions the step needs are driven here through the
// reference classes:
//   * gen_root (src/main.cpp:887-927): bbox loop restated, octree
//     construction and octree::add_particle are the reference's code;
//   * calc_mass / calc_com: the reference's octree methods, unchanged;
//   * the walk (src/main.cpp:335-421) and gravity (:306-333): restated over
//     the reference's vector/octree methods (distance, magnitude, normalize,
//     operator*= ... all from the reference's vector.cpp);
//   * particle::update(dt): the reference's method, unchanged.
// Nothing from the reference's sources is copied into this repository.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <set>
#include <unordered_map>
#include <vector>

#include "octree.h"
#include "particle.h"
#include "vector.h"

namespace {

// src/main.cpp:887-927 (gen_root), bbox and the extent tie-break.
octree* make_root(std::vector<particle*>& ps) {
  datatype min_x = std::numeric_limits<datatype>::max();
  datatype min_y = std::numeric_limits<datatype>::max();
  datatype min_z = std::numeric_limits<datatype>::max();
  datatype max_x = std::numeric_limits<datatype>::lowest();
  datatype max_y = std::numeric_limits<datatype>::lowest();
  datatype max_z = std::numeric_limits<datatype>::lowest();
  for (particle* p : ps) {
    vector t = *(p->get_pos());
    if (t.get_x() < min_x) min_x = t.get_x();
    if (t.get_x() > max_x) max_x = t.get_x();
    if (t.get_y() < min_y) min_y = t.get_y();
    if (t.get_y() > max_y) max_y = t.get_y();
    if (t.get_z() < min_z) min_z = t.get_z();
    if (t.get_z() > max_z) max_z = t.get_z();
  }
  vector origin((min_x + max_x) / 2.0, (min_y + max_y) / 2.0,
                (min_z + max_z) / 2.0);
  datatype size;
  if ((max_x - min_x) > (max_y - min_y) && (max_x - min_x) > (max_z - min_z))
    size = (max_x - min_x) + 2.0;
  else if ((max_y - min_y) > (max_x - min_x) &&
           (max_y - min_y) > (max_z - min_z))
    size = (max_y - min_y) + 2.0;
  else
    size = (max_z - min_z) + 2.0;
  return new octree(&origin, size);
}

// Gravity of one node on one particle, over the reference's vector
// methods (src/main.cpp:306-333).
vector node_gravity(particle* par, octree* node, bool damping) {
  if (par == node->get_particle()) return vector(0, 0, 0);
  vector acc = *(node->get_com());
  acc -= *(par->get_pos());
  datatype r_sq;
  if (!damping)
    r_sq = pow(acc.magnitude(), -2);
  else
    r_sq = 1.0 / (pow(acc.magnitude(), 2) + exp(-1 * acc.magnitude()));
  acc.normalize();
  acc *= node->get_mass();
  acc *= 6.67384e-11;
  acc *= r_sq;
  return acc;
}

struct RefState {
  std::vector<particle*> ps;
  octree* root = nullptr;
  ~RefState() {
    delete root;
    for (particle* p : ps) delete p;
  }
};

void load(RefState& s, size_t n, const double* pos, const double* vel,
          const double* mass) {
  for (size_t i = 0; i < n; i++) {
    vector p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    vector v = vel ? vector(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2])
                   : vector(0, 0, 0);
    vector a(0, 0, 0);
    s.ps.push_back(new particle(&p, &v, &a, mass[i]));
  }
}

// ref_set_threaded(1): build() makes the tree of the threads > 1 path.
int g_threaded = 0;

// gen_root_threaded (src/main.cpp:817-886) with one thread: the root
// allocates its eight children (octree::allocate_child), then each child
// takes, in set order, the particles its octree::inside() accepts
// (octree::add_particle); mass and COM by the reference's
// calc_mass_threaded / calc_com_threaded (src/octree.cpp:230-313).
void build(RefState& s) {
  s.root = make_root(s.ps);
  if (g_threaded) {
    for (unsigned int i = 0; i < 8; i++) s.root->allocate_child(i);
    for (unsigned int i = 0; i < 8; i++) {
      octree* c = s.root->get_child(i);
      for (particle* p : s.ps)
        if (c->inside(p)) c->add_particle(p);
    }
    s.root->calc_mass_threaded();
    s.root->calc_com_threaded();
    return;
  }
  for (particle* p : s.ps) s.root->add_particle(p);
  s.root->calc_mass();
  s.root->calc_com();
}

// BFS walk of src/main.cpp:368-400 for one particle.
void walk(RefState& s, particle* curr, datatype theta, bool damping,
          long* cnt) {
  std::queue<octree*> nodes;
  curr->set_acc_zero();
  nodes.push(s.root);
  long c = 0;
  while (!nodes.empty()) {
    octree* node = nodes.front();
    nodes.pop();
    if (node->get_side() / distance(node->get_com(), curr->get_pos()) >
            theta &&
        node->get_particle() == NULL) {
      for (unsigned int i = 0; i < 8; i++)
        if (node->get_child(i) != NULL) nodes.push(node->get_child(i));
    } else if (curr != node->get_particle()) {
      vector g = node_gravity(curr, node, damping);
      curr->set_acc_offset(&g);
      c++;
    }
  }
  if (cnt) *cnt = c;
}

// collide() of src/main.cpp:423-430 (main.cpp does not compile here):
// weighted_average and the particle constructor are the reference's code.
particle* ref_collide(particle* a, particle* b) {
  vector pos = weighted_average(a->get_pos(), b->get_pos(), a->get_mass(), b->get_mass());
  vector vel = weighted_average(a->get_vel(), b->get_vel(), a->get_mass(), b->get_mass());
  vector null(0, 0, 0);
  return new particle(&pos, &vel, &null, a->get_mass() + b->get_mass());
}

// The walk of src/main.cpp:368-400 with the collision test of :386-396:
// an accepted leaf whose particle is closer than `range` (the reference's
// distance()) records the pair, lower-x particle first, as indices.
void walk_collide(RefState& s, particle* curr, datatype theta, bool damping, double range,
                  const std::unordered_map<particle*, size_t>& index,
                  std::set<std::pair<size_t, size_t>>& pairs) {
  std::queue<octree*> nodes;
  curr->set_acc_zero();
  nodes.push(s.root);
  while (!nodes.empty()) {
    octree* node = nodes.front();
    nodes.pop();
    if (node->get_side() / distance(node->get_com(), curr->get_pos()) > theta &&
        node->get_particle() == NULL) {
      for (unsigned int i = 0; i < 8; i++)
        if (node->get_child(i) != NULL) nodes.push(node->get_child(i));
    } else if (curr != node->get_particle()) {
      particle* o = node->get_particle();
      if (o != NULL && distance(o->get_pos(), curr->get_pos()) < range) {
        const size_t a = index.at(o), b = index.at(curr);
        if (o->get_pos()->get_x() < curr->get_pos()->get_x())
          pairs.insert({a, b});
        else
          pairs.insert({b, a});
      }
      vector g = node_gravity(curr, node, damping);
      curr->set_acc_offset(&g);
    }
  }
}

}  // namespace

extern "C" {

int ref_datatype_bytes(void) { return (int)sizeof(datatype); }

// The bytes src/main.cpp:737 writes per particle: the raw object.
void ref_particle_raw(size_t n, const double* pos, const double* vel,
                      const double* acc, const double* mass,
                      unsigned char* out) {
  for (size_t i = 0; i < n; i++) {
    vector p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    vector v(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    vector a(acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]);
    particle par(&p, &v, &a, mass[i]);
    std::memcpy(out + i * sizeof(particle), (const void*)&par,
                sizeof(particle));
  }
}
int ref_particle_bytes(void) { return (int)sizeof(particle); }

// |a| and |v| of each particle from the reference's own adaptive-dt hook,
// particle::update(dv, dp) (src/particle.cpp:76-80): dv = |acceleration|,
// dp = |velocity|, both by vector::magnitude.
void ref_update_mags(size_t n, const double* vel, const double* acc, double* dv,
                     double* dp) {
  vector p(0, 0, 0);
  for (size_t i = 0; i < n; i++) {
    vector v(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    vector a(acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]);
    particle par(&p, &v, &a, 1);
    datatype x = 0, y = 0;
    par.update(x, y);
    dv[i] = x;
    dp[i] = y;
  }
}
int ref_octree_bytes(void) { return (int)sizeof(octree); }

void ref_set_threaded(int on) { g_threaded = on; }

// Builds the reference tree and dumps it in pre-order (children 0..7).
// Returns the node count; writes at most `cap` nodes (and at most `cap`
// entries of order).  Per node: depth, leaf particle index (or -1), first /
// count in the DFS leaf order, side, mass, COM; `order` gets the particle
// indices in DFS leaf order.
long ref_tree_dump(size_t n, const double* pos, const double* mass,
                   long cap, int* depth, int* leaf, long* first, long* count,
                   double* side, double* nmass, double* com, int* order) {
  RefState s;
  load(s, n, pos, nullptr, mass);
  build(s);
  std::unordered_map<particle*, int> index_of;
  for (size_t k = 0; k < n; k++) index_of[s.ps[k]] = (int)k;
  struct Item {
    octree* node;
    int d;
    int state;
    long slot;
  };
  std::vector<Item> st;
  st.push_back({s.root, 0, 0, 0});
  long out = 0, nleaf = 0;
  while (!st.empty()) {
    Item it = st.back();
    st.pop_back();
    octree* nd = it.node;
    if (it.state == 0) {
      long o = out++;
      bool w = o < cap;
      particle* p = nd->get_particle();
      int li = p != NULL ? index_of[p] : -1;
      if (w) {
        depth[o] = it.d;
        leaf[o] = li;
        first[o] = nleaf;
        side[o] = nd->get_side();
        nmass[o] = nd->get_mass();
        com[3 * o] = nd->get_com()->get_x();
        com[3 * o + 1] = nd->get_com()->get_y();
        com[3 * o + 2] = nd->get_com()->get_z();
      }
      if (p != NULL) {
        if (nleaf < cap) order[nleaf] = li;
        nleaf++;
        if (w) count[o] = 1;
      } else {
        st.push_back({nd, it.d, 1, o});
        for (int i = 7; i >= 0; i--)
          if (nd->get_child(i) != NULL)
            st.push_back({nd->get_child(i), it.d + 1, 0, 0});
      }
    } else if (it.slot < cap) {
      count[it.slot] = nleaf - first[it.slot];
    }
  }
  return out;
}

// Accelerations by the reference walk for all particles; interactions
// (may be null) gets the number of gravity() evaluations per particle.
void ref_forces(size_t n, const double* pos, const double* mass,
                double theta, int damping, double* acc, long* interactions) {
  RefState s;
  load(s, n, pos, nullptr, mass);
  build(s);
  for (size_t i = 0; i < n; i++) {
    long c = 0;
    walk(s, s.ps[i], (datatype)theta, damping != 0, &c);
    vector* a = s.ps[i]->get_acc();
    acc[3 * i] = a->get_x();
    acc[3 * i + 1] = a->get_y();
    acc[3 * i + 2] = a->get_z();
    if (interactions) interactions[i] = c;
  }
}

// `frames` frames of the serial reference loop (build, mass, COM, walk,
// particle::update), in place.  acc receives the last frame's forces.
void ref_run(size_t n, double* pos, double* vel, double* acc,
             const double* mass, double theta, double dt, int damping,
             int frames) {
  RefState s;
  load(s, n, pos, vel, mass);
  for (int f = 0; f < frames; f++) {
    build(s);
    for (size_t i = 0; i < n; i++)
      walk(s, s.ps[i], (datatype)theta, damping != 0, nullptr);
    delete s.root;
    s.root = nullptr;
    for (size_t i = 0; i < n; i++) s.ps[i]->update((datatype)dt);
  }
  for (size_t i = 0; i < n; i++) {
    particle* p = s.ps[i];
    pos[3 * i] = p->get_pos()->get_x();
    pos[3 * i + 1] = p->get_pos()->get_y();
    pos[3 * i + 2] = p->get_pos()->get_z();
    vel[3 * i] = p->get_vel()->get_x();
    vel[3 * i + 1] = p->get_vel()->get_y();
    vel[3 * i + 2] = p->get_vel()->get_z();
    acc[3 * i] = p->get_acc()->get_x();
    acc[3 * i + 1] = p->get_acc()->get_y();
    acc[3 * i + 2] = p->get_acc()->get_z();
  }
}

// Collision frames (src/main.cpp:1420-1492 with config.collide) over the
// reference's classes: walk_collide, then the merge of src/main.cpp:472-498
// (an update table redirects merged particles to their merger; a pair whose
// ends already share one is skipped) with ref_collide, then update_collision
// (:500-550) and particle::update.  The pairs are merged in increasing
// (first, second) index order: the reference's unordered_set order is not
// reproducible.  Output: survivors in index order, each merged particle in
// its lowest-index member's place.  Returns the new count; pairs_out gets
// the distinct pairs recorded over all frames.
long ref_run_collide(size_t n, double* pos, double* vel, double* acc, double* mass,
                     double theta, double dt, int damping, double range, int frames,
                     long* pairs_out) {
  RefState s;
  load(s, n, pos, vel, mass);
  long npairs = 0;
  for (int f = 0; f < frames; f++) {
    std::unordered_map<particle*, size_t> index;
    for (size_t i = 0; i < s.ps.size(); i++) index[s.ps[i]] = i;
    build(s);
    std::set<std::pair<size_t, size_t>> pairs;
    for (particle* p : s.ps)
      walk_collide(s, p, (datatype)theta, damping != 0, range, index, pairs);
    delete s.root;
    s.root = nullptr;
    npairs += (long)pairs.size();
    std::unordered_map<particle*, particle*> table;
    std::vector<particle*> made;
    auto resolve = [&](particle* p) {
      for (auto it = table.find(p); it != table.end(); it = table.find(p)) p = it->second;
      return p;
    };
    for (const auto& q : pairs) {
      particle* a = resolve(s.ps[q.first]);
      particle* b = resolve(s.ps[q.second]);
      if (a == b) continue;
      particle* m = ref_collide(a, b);
      made.push_back(m);
      table[a] = m;
      table[b] = m;
    }
    std::vector<particle*> next;
    std::unordered_map<particle*, bool> placed;
    for (particle* p : s.ps) {
      if (table.find(p) == table.end()) {
        next.push_back(p);
        continue;
      }
      particle* r = resolve(p);
      if (!placed[r]) {
        placed[r] = true;
        next.push_back(r);
      }
      delete p;
    }
    for (particle* m : made)
      if (!placed[m]) delete m;
    s.ps.swap(next);
    for (particle* p : s.ps) p->update((datatype)dt);
  }
  for (size_t i = 0; i < s.ps.size(); i++) {
    particle* p = s.ps[i];
    pos[3 * i] = p->get_pos()->get_x();
    pos[3 * i + 1] = p->get_pos()->get_y();
    pos[3 * i + 2] = p->get_pos()->get_z();
    vel[3 * i] = p->get_vel()->get_x();
    vel[3 * i + 1] = p->get_vel()->get_y();
    vel[3 * i + 2] = p->get_vel()->get_z();
    acc[3 * i] = p->get_acc()->get_x();
    acc[3 * i + 1] = p->get_acc()->get_y();
    acc[3 * i + 2] = p->get_acc()->get_z();
    mass[i] = p->get_mass();
  }
  if (pairs_out) *pairs_out = npairs;
  return (long)s.ps.size();
}

}  // extern "C"
