// Sim-dir configuration and per-source driver, mirroring fast-wave's
// include/simulation.hpp + src/config_parsing.cpp and big-wave's
// wavesim.run_simulation (big-wave/wavesim.py:84-298).  Where the two
// references differ, big-wave's semantics are followed.
#pragma once
#include <complex>
#include <functional>
#include <string>
#include <vector>

#include "engine.hpp"

namespace rs {

// An optical element: a grating (big-wave/optical_element.py:116-225) or,
// with `sample` set, a sample grid (big-wave/optical_element.py:380-476:
// grid[z][x] indexes db_list, 0 = vacuum; one propagation per grid row).
struct GratingCfg {
  double pitch = 0, dc0 = 0.5, dc1 = 0.5, z_start = 0, thickness = 0;
  int nr_steps = 1;
  std::vector<double> x_positions{0.0};
  double substrate_thickness = 0;
  std::complex<double> db_a{0, 0}, db_b{0, 0}, db_s{0, 0};
  // DBA grating (fast-wave/include/optical_element.hpp:50-63): material A
  // where both the (pitch, dc0..dc1) grating and frac(x/pitch1) < dc1 hold;
  // pitch / dc0 / dc1 above are its pitch0 and dc0 pair, dcb its dc1 pair
  bool dba = false;
  double pitch1 = 0, dcb0 = 0, dcb1 = 0;
  bool sample = false;
  double pixel_size_x = 0, pixel_size_z = 0;
  int64_t rows = 0, cols = 0;
  std::vector<int8_t> grid;                  // rows x cols
  std::vector<std::complex<double>> db_list;  // [0] = vacuum
  double total_thickness() const {
    return sample ? pixel_size_z * (double)rows : thickness + substrate_thickness;
  }
};

struct SourceCfg {
  enum Kind { Point, Vector, Plane } kind = Point;
  double x = 0, z = 0;
  std::string input_path;
};

struct SimConfig {
  int64_t N = 0;
  double dx = 0, z_detector = 0, detector_size = 0, px = 0, py = 1;
  int64_t chunk_size = 0;
  bool f64 = true;               // big-wave default dtype c16 (big-wave/config.py:176-177)
  bool save_final_u_vectors = false;
  std::vector<GratingCfg> elements;
  std::vector<double> cutoff_angles;
  double energy = 0, wl = 0;
  SourceCfg source;
};

std::string subdir_name(int source_idx);  // "%08d"

// Parse sim_dir/config.yaml, computed.yaml and %08d/subconfig.yaml
// (fast-wave/src/config_parsing.cpp:281-307, big-wave/multisim.py:70-112).
SimConfig parse_simdir(const std::string &sim_dir, int source_idx);

// Optional recordings of the first pass through the elements (phase step 0).
struct Recording {
  // History (big-wave/history.py, wavesim.py:153-219): with history_dz > 0,
  // intensity rows every history_dz in free space (re-propagated from the
  // segment's start), after every grating slice and before every sample row;
  // a point source contributes analytic rows up to the first element.
  double history_dz = 0;
  // Keypoints (big-wave/wavesim.py:35-58): the field before (sub 0) and
  // after (sub 1) every element, as Keypoint ops with row 2 * element + sub.
  bool keypoints = false;
  // filled by build_member, one element per history entry in push order: its
  // z, its accumulator row (-1 for an analytic entry) and, for an analytic
  // entry, its intensity row (empty otherwise); n_rows counts the
  // accumulator rows used in total
  std::vector<double> hist_z;
  std::vector<int> hist_row;
  std::vector<std::vector<double>> analytic;
  int n_rows = 0;
};

// Build the member program of one source, incl. phase steps.  Returns the
// number of detector rows (phase steps); history rows follow them.
int build_member(const SimConfig &c, Member &m, int out_base, bool save_u, Recording *rec = nullptr);

// Called with z after every grating slice and before every sample row
// (history entries of big-wave/optical_element.py:182-187, 448-453).
using SliceHook = std::function<void(double z)>;

// Slice loop + substrate of one grating as member ops, using
// x_positions[step] (big-wave/optical_element.py:134-206).
void grating_ops(const SimConfig &c, const GratingCfg &g, int step, double fc, std::vector<Op> &ops,
                 const SliceHook &hook = nullptr);

// Row loop of a sample element as member ops (one SAMPLE mask and one
// propagation over pixel_size_z per grid row), its pixel tables appended to
// m.samples (big-wave/optical_element.py:432-476).
void sample_ops(const SimConfig &c, const GratingCfg &g, int step, double fc, Member &m,
                const SliceHook &hook = nullptr);

// detector pixel x positions (k - n // 2) * px (big-wave/util.py:78-82)
std::vector<double> detector_x(double size, double px);

// Check element ordering / phase-step consistency like big-wave's
// check_simulation_inputs (big-wave/multisim.py:195-242); throws Error.
void check_config(const SimConfig &c);

}  // namespace rs
