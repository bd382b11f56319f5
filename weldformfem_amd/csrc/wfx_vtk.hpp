// One legacy VTK frame of the step's state, on the host in the reference's
// layout (nodal vectors [N][3], element tensors [E][6] xx yy zz xy yz xz,
// connectivity [E][n]); written by wfx_io.cpp for both wfx_write_vtk and the
// asynchronous frames of wfx_frames.hip.
#pragma once

#include <string>
#include <vector>

namespace wfx {

struct VtkFrame {
  int N = 0, E = 0, n = 0;
  std::vector<double> x, u, v, a, mdiag, p, pl_strain, sigma_y, vol, sigma;
  std::vector<int> elnod;
};

// 0 on success, else a message in err
int write_vtk_frame(const char *path, const VtkFrame &fr, std::string &err);

}  // namespace wfx
