/*
 * vgi_image.h — texture images for the glTF importer: PNG decoding to RGBA8
 * and the mip chain of violet's importer.
 *
 * violet decodes glTF images with the `image` crate 0.24 (features jpeg,
 * png; crates/violet/Cargo.toml:21) through the `gltf` crate 1.x, keeps
 * 8-bit RGB and RGBA images (model.rs:816-846) and builds its mips by
 * repeated halving with image::imageops::resize(.., Lanczos3)
 * (model.rs:853-870).  This file restates those steps:
 *   - PNG (ISO/IEC 15948, RFC 1950/1951 for the zlib stream): every colour
 *     type and bit depth, palettes, tRNS and Adam7 interlacing, expanded to
 *     RGBA8 (16-bit samples round to 8 bits, grey replicates to RGB;
 *     violet itself stops on images other than 8-bit RGB or RGBA);
 *   - baseline JPEG (ITU-T T.81 sequential Huffman, JFIF YCbCr): a float
 *     8x8 inverse DCT, the triangle ("fancy") chroma upsampling for 2:1
 *     factors that libjpeg and the jpeg-decoder crate use, and the JFIF
 *     YCbCr -> RGB equations, each rounded to nearest (decoders differ
 *     from one another by a unit or two here; violet decodes with the
 *     jpeg-decoder crate);
 *   - Lanczos3 halving as the image crate 0.24's separable resampler
 *     computes it (vertical pass into f32, then horizontal pass with
 *     round-to-nearest and clamp to [0, 255]; the filter support widened by
 *     the scale ratio; weights normalised per output pixel).
 */
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace vgi {

struct Rgba8Image {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> texels;  // width * height * 4
};

// Decodes a PNG file image; false with a reason in `err` on malformed or
// truncated input, or on sizes beyond max_texels.
bool decode_png(const uint8_t* data, size_t size, Rgba8Image& out, std::string& err,
                uint64_t max_texels = 1ull << 28);

// True if the bytes start with the PNG signature.
bool is_png(const uint8_t* data, size_t size);

// Decodes a baseline (sequential, Huffman) JPEG: greyscale or YCbCr, any
// sampling factors up to 4, restart intervals.  Returns 1 on success, 0 for
// a valid JPEG this decoder does not handle (progressive, arithmetic-coded,
// 12-bit, CMYK), -1 for malformed or truncated input (reason in `err`).
int decode_jpeg(const uint8_t* data, size_t size, Rgba8Image& out, std::string& err,
                uint64_t max_texels = 1ull << 28);

// True if the bytes start with a JPEG SOI marker.
bool is_jpeg(const uint8_t* data, size_t size);

// image::imageops::resize(src, nw, nh, FilterType::Lanczos3).
Rgba8Image resize_lanczos3(const Rgba8Image& src, uint32_t nw, uint32_t nh);

// violet's mip chain: 1 + log2(next_pow2(max(w, h))) levels, level l+1 the
// Lanczos3 resize of level l to max(1, w/2) x max(1, h/2).
std::vector<Rgba8Image> mip_chain(const Rgba8Image& base);

}  // namespace vgi
