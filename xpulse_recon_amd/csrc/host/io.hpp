// File formats of the reference's data path: FP32 grayscale TIFF (TinyTIFF,
// dataset/data_loader.hpp:220-238) and zstd-compressed float spill files
// (data_loader.hpp:239-262), plus the directory scan (dataset.hpp:214-269).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace xpr {

struct TiffInfo {
    int64_t width = 0, height = 0;
    int bits = 0, sample_format = 0, compression = 0, samples = 0;
};

// Header fields of a TIFF; throws std::runtime_error on unreadable files.
TiffInfo tiff_info(const std::string &path);
// Reads an uncompressed single-channel FP32 TIFF.  Throws on other formats.
std::vector<float> tiff_read_f32(const std::string &path, int64_t *width = nullptr,
                                 int64_t *height = nullptr);
void tiff_write_f32(const std::string &path, const float *data, int64_t width, int64_t height);

// zstd level-1 frames of raw floats (the reference's temp files).  When
// libzstd.so.1 cannot be loaded the bytes are stored raw under the same name.
bool zstd_available();
void spill_write(const std::string &path, const float *data, size_t n);
void spill_read(const std::string &path, float *data, size_t n);

// *.tif / *.tiff regular files of a directory, sorted by file name.
std::vector<std::string> list_tiffs(const std::string &dir);

// "<dir>/<prefix>_<index zero-padded to width>.<ext>" (data_loader.hpp:183-199)
std::string numbered_path(const std::string &dir, const std::string &prefix, int64_t index,
                          int width, const std::string &ext);

} // namespace xpr
