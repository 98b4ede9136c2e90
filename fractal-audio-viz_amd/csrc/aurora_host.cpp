// aurora_host.cpp -- the ordinary (Aurora) palette, built on the host with
// the reference's double arithmetic and the host libm's cos / sin
// (aurora_palette_for and colour_basis_for, renderer.cpp:4788-4923).
#include <algorithm>
#include <cmath>

#include "aurora_core.h"

namespace fav {
namespace aurora {

void build_palette(int max_iter, double phase, double vocal, double pitch, const uint8_t* accents,
                   std::vector<uint8_t>& rgb) {
    const int size = std::min(65536, std::max(4096, max_iter * 4));
    const double denominator = static_cast<double>(std::max(1, size - 1));
    const double vocal_mix = std::clamp(vocal, 0.0, 1.0);
    const double split = 5.0 * vocal_mix * vocal_mix;
    const double red_cos = std::cos(phase), red_sin = std::sin(phase);
    const double green_cos = std::cos(phase + split * 0.4), green_sin = std::sin(phase + split * 0.4);
    const double blue_cos = std::cos(phase + split), blue_sin = std::sin(phase + split);
    const double hue_angle = pitch_hue_angle(pitch);
    const bool rotate = std::abs(hue_angle) > 1.0e-15;
    const double hue_cos = rotate ? std::cos(hue_angle) : 1.0;
    const double hue_sin = rotate ? std::sin(hue_angle) : 0.0;
    rgb.resize(static_cast<size_t>(size) * 3);
    for (int index = 0; index < size; ++index) {
        const double angle = 0.15 * static_cast<double>(max_iter) * static_cast<double>(index) / denominator;
        const double cosine = static_cast<double>(static_cast<float>(std::cos(angle)));
        const double sine = static_cast<double>(static_cast<float>(std::sin(angle)));
        double r = (0.5 - 0.5 * (cosine * red_cos + sine * red_sin)) * 140.0;
        double g = (0.5 - 0.5 * (cosine * green_cos + sine * green_sin)) * 140.0;
        double b = (0.5 - 0.5 * (cosine * blue_cos + sine * blue_sin)) * 140.0;
        if (accents) {  // the three waves projected through three RGB accents
            const double rw = std::clamp(static_cast<double>(colour_byte(r)) / 140.0, 0.0, 1.0);
            const double gw = std::clamp(static_cast<double>(colour_byte(g)) / 140.0, 0.0, 1.0);
            const double bw = std::clamp(static_cast<double>(colour_byte(b)) / 140.0, 0.0, 1.0);
            const double a[9] = {static_cast<double>(accents[0]), static_cast<double>(accents[1]),
                                 static_cast<double>(accents[2]), static_cast<double>(accents[3]),
                                 static_cast<double>(accents[4]), static_cast<double>(accents[5]),
                                 static_cast<double>(accents[6]), static_cast<double>(accents[7]),
                                 static_cast<double>(accents[8])};
            r = (rw * a[0] + gw * a[3] + bw * a[6]) / 1.8;
            g = (rw * a[1] + gw * a[4] + bw * a[7]) / 1.8;
            b = (rw * a[2] + gw * a[5] + bw * a[8]) / 1.8;
        }
        if (rotate) rotate_hue(r, g, b, hue_cos, hue_sin);
        rgb[static_cast<size_t>(index) * 3] = colour_byte(r);
        rgb[static_cast<size_t>(index) * 3 + 1] = colour_byte(g);
        rgb[static_cast<size_t>(index) * 3 + 2] = colour_byte(b);
    }
}

}  // namespace aurora
}  // namespace fav
