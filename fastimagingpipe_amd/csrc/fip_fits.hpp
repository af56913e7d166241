// FITS writing shared by the library's own tools (not part of the C ABI).
#pragma once

namespace fip {

// One BITPIX -32 IMAGE extension: naxes[0] is NAXIS1 (fastest axis).
struct FitsImage {
    const char* extname;
    int naxis;
    long naxes[3];
    const float* data;
};

// An empty primary HDU (EXTEND = T) followed by the n extensions, replacing
// an existing file.  Returns a fip_status.
int fits_write_extensions(const char* path, const FitsImage* ext, int n);

}  // namespace fip
