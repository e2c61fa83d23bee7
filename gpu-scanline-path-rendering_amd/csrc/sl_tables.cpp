// sl_tables.cpp -- host-built lookup tables uploaded once per handle:
//  * sample positions (vg_sample_position, ras_qm_mask.cu:118-175; 16 spp extends the
//    8-spp bit-reversal rule to 16 positions);
//  * the 256x256 line -> sample-mask table (QM_Mask_2x2::makeTable, ras_qm_mask.h:432-484),
//    built on the host in IEEE single precision so that every device sees identical bits;
//  * colour decode tables and encode thresholds for the compositor's 8-bit framebuffer.
#include "sl_tables.h"

#include <cmath>
#include <cstring>

namespace sl {

static const double kX32[32] = {
    -0.49479166651144624000, 0.02604168653488159200, -0.24479165673255920000, 0.27604168653488159000,
    -0.36979167163372040000, 0.15104168653488159000, -0.09895834326744079600, 0.38020831346511841000,
    -0.43229166418313980000, 0.06770831346511840800, -0.18229165673255920000, 0.33854168653488159000,
    -0.30729167163372040000, 0.21354168653488159000, -0.05729165673255920400, 0.46354168653488159000,
    -0.46354166790843010000, 0.05729168653488159200, -0.19270834326744080000, 0.30729168653488159000,
    -0.31770832836627960000, 0.16145831346511841000, -0.06770834326744079600, 0.43229168653488159000,
    -0.40104166418313980000, 0.11979168653488159000, -0.13020834326744080000, 0.36979168653488159000,
    -0.27604167163372040000, 0.24479168653488159000, -0.02604165673255920400, 0.47395831346511841000,
};  // vg_32_x (ras_qm_mask.cu:111-113)

void sample_positions(int n, float *xy) {
    if (n == 32) {
        for (int i = 0; i < 32; ++i) {
            xy[2 * i] = (float)(kX32[i] + 0.5);
            xy[2 * i + 1] = (float)((-0.484375 + 0.03125 * i) + 0.5);  // vg_32_y + 0.5
        }
        return;
    }
    const double gap = 1.0 / n, hgap = gap / 2;
    for (int i = 0; i < n; ++i) {
        int bits = 0, v = i, rev = 0;
        while (v) { rev = (rev << 1) | (v & 1); v >>= 1; ++bits; }
        double f = rev / (float)(1 << bits);
        int col = (int)(f * n);
        xy[2 * i] = (float)(col * gap + hgap);
        xy[2 * i + 1] = (float)(i * gap + hgap);
    }
}

void build_pixel_table(int n, std::vector<uint32_t> &out) {
    const int nw = n == 8 ? 1 : (n == 16 ? 2 : 4);
    float s[64];
    sample_positions(n, s);
    // cell coordinates of every (pixel q = px*2+py, sample i)
    float cx[4][32], cy[4][32];
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < n; ++i) {
            cx[q][i] = (s[2 * i] + (float)(q >> 1)) * .5f;
            cy[q][i] = (s[2 * i + 1] + (float)(q & 1)) * .5f;
        }
    out.assign((size_t)65536 * nw, 0u);
    for (int y = 0; y < 256; ++y)
        for (int x = 0; x < 256; ++x) {
            const float xf = ((float)x + 0.5f) * (1.f / 256.f);
            const float yf = ((float)y + 0.5f) * (1.f / 256.f);
            const float ax = 2.f * (xf - 0.5f), ay = 2.f * (yf - 0.5f);
            const float len = std::sqrt(ax * ax + ay * ay);
            const float inv = 1.f / len;
            float nx = inv * ax, ny = inv * ay;
            float c = ((1.f - len) > 0.f ? (1.f - len) : 0.f) * (1.f / 0.5f);
            if (nx < 0.f) { nx = -nx; ny = -ny; c = -c; }
            if (x == 37 && y == 218) {
                const float il = 1.f / std::sqrt(1.f * 1.f + (-1.f) * (-1.f));
                nx = 1.f * il;
                ny = -1.f * il;
                c = 0.f;
            }
            c += 0.5f * (nx + ny);
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            for (int q = 0; q < 4; ++q)
                for (int i = 0; i < n; ++i) {
                    if ((nx * cx[q][i] + ny * cy[q][i]) - c > 0.f) {
                        if (n == 8) w[0] |= 1u << (q * 8 + i);
                        else if (n == 16) w[q >> 1] |= 1u << ((q & 1) * 16 + i);
                        else w[q] |= 1u << i;
                    }
                }
            for (int k = 0; k < nw; ++k) out[(size_t)(y * 256 + x) * nw + k] = w[k];
        }
}

static double srgb_to_linear(double s) { return s <= 0.04045 ? s / 12.92 : std::pow((s + 0.055) / 1.055, 2.4); }

static float smallest_float_at_least(double d) {
    float f = (float)d;
    while ((double)f < d) f = std::nextafter(f, INFINITY);
    for (;;) {
        float g = std::nextafter(f, -INFINITY);
        if ((double)g >= d) f = g;
        else break;
    }
    return f;
}

bool build_color_tables(int srgb, ColorTables &t) {
    for (int k = 0; k < 256; ++k) {
        t.dec[k] = srgb ? (float)srgb_to_linear(k / 255.0) : (float)k / 255.f;
        if (k == 0) {
            t.thr[0] = -INFINITY;
        } else {
            double mid = (k - 0.5) / 255.0;
            t.thr[k] = smallest_float_at_least(srgb ? srgb_to_linear(mid) : mid);
        }
    }
    // lut[i] = code of i/4096; valid only if no bucket (i/4096, (i+1)/4096) holds more
    // than one threshold, which the check below enforces.
    int k = 0;
    for (int i = 0; i < 4096; ++i) {
        const float lo = (float)i / 4096.f;
        while (k < 255 && lo >= t.thr[k + 1]) ++k;
        t.lut[i] = (uint8_t)k;
        const float hi = (float)(i + 1) / 4096.f;
        int inside = 0;
        for (int j = k + 1; j < 256; ++j) {
            if (t.thr[j] > lo && t.thr[j] < hi) ++inside;
            if (t.thr[j] >= hi) break;
        }
        if (inside > 1) return false;
    }
    return true;
}

}  // namespace sl
