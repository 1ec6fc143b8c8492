// Element stream: the mesh as the reference's gmsh reader hands it to the
// partitioner (RootStructure.cpp:428-487): vertices, then elements with a
// gmsh type code, a physical tag and 0-based vertex ids.  Synthetic cases are
// generated straight into this form so that the ingest path (element-order
// permutation, face builder, geometry, CSR) is exercised exactly as it would
// be for a mesh file.
#pragma once
#include <cstdint>
#include <vector>

namespace s355 {

// gmsh element type codes used by the reference (geometry.cpp:50, 136-201)
enum GmshType : int { kTri = 2, kQuad = 3, kTet = 4, kHex = 5, kPrism = 6, kPyramid = 7 };

inline int gmsh_nodes(int type) {
  switch (type) {
    case kTri: return 3;
    case kQuad: return 4;
    case kTet: return 4;
    case kHex: return 8;
    case kPrism: return 6;
    case kPyramid: return 5;
    default: return -1;
  }
}

// Boundary-region semantics of BdRegion (BasicType.h:247-310):
// type1 1 wall, 2 inlet, 3 outlet, 4 symmetry, 5 volume; init values are
// [u,v,w,rho,T,mu,k,eps] for volumes and [u,v,w,p,rho,T,k,eps] for inlets.
struct Region {
  int rid = 0;
  int type1 = -1;
  int type2 = 0;
  double fixed_value = 0.0;
  double init[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

struct ElementStream {
  std::vector<double> xyz;       // 3 * n_vert
  std::vector<int> type;         // gmsh type per element
  std::vector<int> tag;          // physical tag (region id) per element
  std::vector<int> vptr{0};      // CSR offsets into vidx, n_elem + 1
  std::vector<int> vidx;         // 0-based vertex ids
  std::vector<Region> regions;   // region table (rid 0 = default fluid)

  int n_vert() const { return static_cast<int>(xyz.size() / 3); }
  int n_elem() const { return static_cast<int>(type.size()); }
  void add(int t, int tg, const int* v) {
    type.push_back(t);
    tag.push_back(tg);
    for (int k = 0; k < gmsh_nodes(t); ++k) vidx.push_back(v[k]);
    vptr.push_back(static_cast<int>(vidx.size()));
  }
};

// Deterministic, platform-independent generator (splitmix64).
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double sym() { return 2.0 * uniform() - 1.0; }                              // [-1,1)
};

// ---- generators (case_gen.cpp) -------------------------------------------

struct BoxSpec {
  int nx = 8, ny = 8, nz = 8;
  double lx = 1, ly = 1, lz = 1;
  double x0 = 0, y0 = 0, z0 = 0;
  double jitter = 0.0;      // node perturbation as a fraction of the local spacing
  double split_frac = 0.0;  // fraction of hexes split into 6 tets (Kuhn split)
  uint64_t seed = 0;
  // physical tags of the six sides: x-, x+, y-, y+, z-, z+ ; and the volume
  int side_tag[6] = {1, 1, 1, 1, 1, 1};
  int volume_tag = 100;
  bool all_tets = false;    // split every hex (tetrahedral mesh)
};

ElementStream make_box(const BoxSpec& spec);

// Backward-facing step: inlet channel of height h on top of the step,
// expansion ratio 2.  Cells with x < 0 and y < h are removed.
struct StepSpec {
  double h = 1.0;
  double upstream = 4.0, downstream = 30.0, span = 4.0;
  int n_per_h = 4;          // cells per step height
  double jitter = 0.10;
  uint64_t seed = 0;
  bool tets = true;
  int tag_inlet = 2, tag_outlet = 3, tag_wall = 1, tag_sym = 4, volume_tag = 100;
};

ElementStream make_step(const StepSpec& spec);

}  // namespace s355
