// flip_view.h -- internal access to a FLIP handle's device state for the U-Net
// warm start (both live in libaqflip).
#pragma once

#include <string>

#include "../../include/aqflip.h"
#include "aqflip_internal.h"

namespace aqf {

struct FlipView {
  Geom G;
  Grid* gr;
  int* pb;  // current pressure buffer index
  hipStream_t stream;
};

FlipView flip_view(aq_flip* h);
void set_last_error(const std::string& msg);

}  // namespace aqf
