// es_png.hpp — the images of output_type "all" (Logger::LogImages, src/Logger.cpp:278-334).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace es_host {

// Logger::WriteImageSection (Logger.cpp:278-318): panel `image_id` (0, 1, 2) of
// a (3 * dim) x dim RGB image; grey (v - min) / max * 255, or, `colored`,
// red c / max * 255 for c >= 0 and blue c / min * 255 for c < 0.
void write_image_section(std::vector<uint8_t>& pixels, const double* values, int dim, int image_id, bool colored);

// Logger::LogImages (Logger.cpp:320-334): lifetimes | sigma_xx | sigma_xy (coloured)
// side by side, written as an 8-bit RGB PNG.  Returns false if the file cannot be written.
bool log_images(const std::string& path, int dim, const double* lifetimes, const double* sigma_xx,
                const double* sigma_xy);

// An 8-bit RGB PNG of width x height from row-major pixels (3 bytes each).
bool write_png_rgb(const std::string& path, int width, int height, const std::vector<uint8_t>& pixels);

}  // namespace es_host
