// Host-side pieces shared by the C ABI files: the result object behind
// vltr_result*, the thread-local error setter and a tessellation call that
// can also hand back each contact's restricted contour.
#pragma once

#include <string>
#include <vector>

#include "group_clip.h"
#include "voronota_ltr.h"

struct vltr_result {
  size_t num_balls = 0;
  std::vector<vltr_contact> contacts;
  std::vector<vltr_cell> cells;
  bool has_network = false;
  std::vector<vltr_cell_vertex> vertices;
  std::vector<vltr_cell_edge> edges;
};

namespace vltr {

// Sets vltr_last_error's message and returns code.
int host_fail(int code, const std::string& msg);

// compute_tessellation on host balls (the vltr_compute_tessellation path);
// with with_vertices and nets != nullptr also returns the contours, one per
// contact in contact order.
int host_tessellate(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe, const vltr_periodic_box* pbox,
                    const int32_t* groups, size_t n_groups, int with_vertices, vltr_result** out,
                    std::vector<ContactNet>* nets = nullptr);

}  // namespace vltr
