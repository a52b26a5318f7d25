// rd_image.h — image files on either side of the path: the decoders and
// encoders behind TextureUtility.LoadImageFromPath / SavePng / SaveOpenExr
// (Radiantium.Core/TextureUtility.cs:54-203), written for this library.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace rd {

struct ImageError : std::runtime_error {
    int code;   // RD_ERR_*
    ImageError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// A decoded file: rows top first, `channels` floats per pixel, values as
// Magick.NET-Q16-HDRI hands them to LoadImageFromPath after its division by
// Quantum.Max (so 8-bit samples are v / 255 and float samples pass through
// the float quantum: (float)(65535 * v) / 65535).
struct RawImage {
    int w = 0, h = 0, channels = 0;
    std::vector<float> px;   // (y * w + x) * channels
};

// Decodes PNG, OpenEXR (scanline; NONE, RLE, ZIPS, ZIP, PIZ; HALF / FLOAT /
// UINT), Radiance RGBE (.hdr) and PFM, chosen by the file's magic bytes.
// Throws ImageError.
RawImage decode_image(const std::string& bytes, const std::string& name);

// IEEE half <-> float, round to nearest even.
float half_to_float(uint16_t h);
uint16_t float_to_half(float f);

}  // namespace rd
