// ConvertNII -- per-voxel result vectors (.nm) back to NIfTI volumes through
// the mask (drop-in for U/ConvertNII/ConvertNII/ConvertNII.cpp:16-179).
//
//   ConvertNII Directory_containing_files_to_be_converted niiFile_for_mask [threshold_for_mask]
//
// Every <dir>/*.nm (int32 N, float32[N]) becomes <dir>/<name>.nii: the mask's
// 352-byte header with datatype 16 (float32) and bitpix 32, then one float per
// mask voxel, the .nm values in voxel order at voxels with mask >= threshold
// (default 0.5) and 0 elsewhere.  A file whose N differs from the mask's voxel
// count is reported (stdout and BNA_time_log) and skipped.  Host-only: no GPU
// work is involved.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "io.hpp"

int main(int argc, char **argv)
{
    if (argc < 3 || argc > 4) {
        std::cerr << "Input format: ConvertNII Directory_containing_files_to_be_converted niiFile_for_mask "
                     "[ |threshold_for_mask]\nFor example: ConvertNII X_deg.nm mask.nii [ |0.4]" << std::endl;
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && (dir.back() == '/' || dir.back() == '\\')) dir.pop_back();
    const auto files = pgio::list_dir(dir, ".nm");
    std::cout << files.size() << " files to be processed." << std::endl;

    std::vector<float> mask;
    int dims[3];
    short hdr[176];
    std::string err;
    if (!pgio::read_mask(argv[2], mask, dims, err, hdr)) {
        std::cerr << err << std::endl;
        return 2;
    }
    // a double cut, compared with each float mask value promoted to double
    // (ConvertNII.cpp:13, :107-110; CUCorMat compares in float instead)
    const double cut = argc == 4 ? std::atof(argv[3]) : 0.5;
    long long nvox = 0;
    for (float m : mask)
        if ((double)m >= cut) nvox++;
    hdr[35] = 16;  // datatype float32
    hdr[36] = 32;  // bitpix
    std::vector<float> vol(mask.size());
    for (const auto &name : files) {
        const std::string a = pgio::join(dir, name);
        std::cout << "\nconvert analysis for " << a << " ..." << std::endl;
        std::ifstream fi(a, std::ios::binary);
        int32_t n = -1;
        fi.read(reinterpret_cast<char *>(&n), 4);
        if (!fi.good()) {
            std::cerr << "Can't open\t" << a << std::endl;
            return 2;
        }
        if ((long long)n != nvox) {
            std::cout << "Mask file doesn't match to the file " << name << "\nContinue to convert the next file"
                      << std::endl;
            pgio::time_log("Mask file doesn't match to the file " + name);
            continue;
        }
        std::vector<float> x((size_t)n);
        fi.read(reinterpret_cast<char *>(x.data()), (std::streamsize)n * 4);
        if (!fi.good()) {
            std::cerr << "truncated " << a << std::endl;
            return 2;
        }
        size_t l = 0;
        for (size_t k = 0; k < mask.size(); k++) vol[k] = (double)mask[k] >= cut ? x[l++] : 0.0f;
        const std::string out = pgio::strip_ext(a) + ".nii";
        std::ofstream fo(out, std::ios::binary);
        fo.write(reinterpret_cast<const char *>(hdr), sizeof hdr);
        fo.write(reinterpret_cast<const char *>(vol.data()), (std::streamsize)(vol.size() * sizeof(float)));
        if (!fo.good()) {
            std::cerr << "can't write " << out << std::endl;
            return 2;
        }
        std::cout << "Convert to .nii completed!" << std::endl;
    }
    std::cout << "===========================================================" << std::endl;
    return 0;
}
