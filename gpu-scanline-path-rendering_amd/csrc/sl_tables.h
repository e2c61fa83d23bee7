// sl_tables.h -- host-built lookup tables of the renderer.
#pragma once

#include <stdint.h>
#include <vector>

namespace sl {

// Sample positions of one pixel for n = 8, 16 or 32 samples, x/y pairs in [0,1)^2.
void sample_positions(int n, float *xy);

// 256x256 line -> sample-mask table, words per texel: 1 (8 spp), 2 (16 spp), 4 (32 spp).
void build_pixel_table(int n, std::vector<uint32_t> &out);

struct ColorTables {
    float dec[256];    // 8-bit code -> value (sRGB decoded, or k/255)
    float thr[256];    // smallest value encoding to k (thr[0] = -inf)
    uint8_t lut[4096]; // code of i/4096, refined by one threshold compare
};
bool build_color_tables(int srgb, ColorTables &t);

}  // namespace sl
