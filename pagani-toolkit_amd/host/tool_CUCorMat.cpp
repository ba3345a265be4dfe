// CUCorMat -- voxel time series to thresholded CSR networks (drop-in for
// U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:51-378 and the weighted twin
// W/CorMat/CUCorMat/CUCorMat/CorMat.cpp).
//
//   CUCorMat [--weighted] [--fp32] Dir [Mask_Path] maskthr avg cormat {r|s} thr...
//
// Source form (mask.nii inside Dir, outputs beside the inputs) and the
// documented form with Mask_Path (outputs into Dir/unweighted or
// Dir/weighted), as SP/script.bat uses it.  Output per subject and threshold:
//   <subject>_<N>_spa%.3f%%_cor%.3f.csr   (weighted: ..._weighted.csr)
// avg must be 'n' (per-subject networks); cormat must be 'n'.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "tool_common.hpp"

int main(int argc, char *argv[])
{
    bool weighted = false;
    int prec = PG_PREC_BF16X3;
    int a0 = 1;
    while (a0 < argc && std::strncmp(argv[a0], "--", 2) == 0) {
        if (!std::strcmp(argv[a0], "--weighted")) weighted = true;
        else if (!std::strcmp(argv[a0], "--fp32")) prec = PG_PREC_FP32;
        else if (!std::strcmp(argv[a0], "--bf16x3")) prec = PG_PREC_BF16X3;
        else {
            std::cerr << "unknown option " << argv[a0] << "\n";
            return 1;
        }
        a0++;
    }
    if (argc - a0 < 5) {
        std::cerr << "Input format: CUCorMat [--weighted] [--fp32] Dir_for_BOLD [Mask_Path] threshold_for_mask(0~1) "
                     "to_average(n) to_save_cormatrix(n) threshold_type(r/s) threshold(s)\n"
                     "For example: CUCorMat /data/BOLD 0.5 n n r 0.2 0.25 0.3\n";
        return 1;
    }
    const std::string dir = pgtool::dir_arg(argv[a0]);
    std::string mask_path;
    int k = a0 + 1;
    const std::string maybe = argv[k];
    const bool documented = maybe.size() > 4 && maybe.substr(maybe.size() - 4) == ".nii";
    if (documented) {
        mask_path = maybe;
        k++;
    } else {
        mask_path = pgio::join(dir, "mask.nii");
    }
    if (argc - k < 4) {
        std::cerr << "missing arguments\n";
        return 1;
    }
    const float prob_cut = (float)std::atof(argv[k]);
    const char avg = argv[k + 1][0];
    const char save = argv[k + 2][0];
    const char kind = argv[k + 3][0];
    if (avg != 'n' && avg != 'N') {
        std::cerr << "to_average: only 'n' (per-subject networks) is supported\n";
        return 1;
    }
    if (save != 'n' && save != 'N') {
        std::cerr << "to_save_cormatrix: only 'n' is supported (the dense matrix is never formed)\n";
        return 1;
    }
    bool rmode;
    if (kind == 'r' || kind == 'R') rmode = true;
    else if (kind == 's' || kind == 'S') rmode = false;
    else {
        std::cerr << "threshold type error!\nr for correlation threshold\ns for spasity threshold\n";
        return 1;
    }
    std::vector<float> thr;
    for (int q = k + 4; q < argc; q++) thr.push_back((float)std::atof(argv[q]));
    if (thr.empty()) {
        std::cerr << "no threshold given\n";
        return 1;
    }
    // U CorMat.cpp:175-181: the smallest threshold is swapped to the front
    if (rmode) {
        size_t mi = 0;
        for (size_t q = 0; q < thr.size(); q++)
            if (thr[q] < thr[mi]) mi = q;
        std::swap(thr[0], thr[mi]);
    }

    std::vector<float> mask;
    int dims[3];
    std::string err;
    if (!pgio::read_mask(mask_path, mask, dims, err)) {
        std::cerr << err << "\n";
        return 2;
    }
    std::string outdir = dir;
    if (documented) {
        outdir = pgio::join(dir, weighted ? "weighted" : "unweighted");
        pgio::make_dir(outdir);
    }
    const auto files = pgio::list_dir(dir, ".nii", "mask.nii");
    std::cout << files.size() << " file(s) to be processed." << std::endl;
    pgtool::Ctx ctx;
    pgtool::Timer total;
    for (const auto &fname : files) {
        const std::string a = pgio::join(dir, fname);
        if (documented && a == mask_path) continue;
        std::cout << "\ncalculating correlation for " << a << " ..." << std::endl;
        int n = 0, t = 0;
        std::vector<float> bold;
        if (!pgio::read_bold(a, mask, prob_cut, n, t, bold, err)) {
            std::cerr << err << "\n";
            return 2;
        }
        std::cout << "Data size: " << dims[0] << "x" << dims[1] << "x" << dims[2] << ", Grey voxel count: " << n
                  << ", BOLD length: " << t << "." << std::endl;
        const std::string base = pgio::join(outdir, pgio::strip_ext(fname)) + "_" + std::to_string(n);
        for (float th : thr) {
            pg_corr_opts o;
            o.thr_kind = rmode ? PG_THR_R : PG_THR_SPARSITY;
            o.thr = th;
            o.weighted = weighted ? 1 : 0;
            o.prec = prec;
            o.layout = PG_BOLD_TIME_MAJOR;
            pg_corr_info info;
            pgtool::Timer tm;
            ctx.check(pg_corr_build(ctx.c, bold.data(), n, t, &o, &info), "pg_corr_build");
            std::vector<int32_t> row(n + 1), col((size_t)std::max<int64_t>(info.nnz, 1));
            std::vector<float> val(weighted ? col.size() : 0);
            ctx.check(pg_corr_csr(ctx.c, row.data(), col.data(), weighted ? val.data() : nullptr), "pg_corr_csr");
            const double secs = tm.seconds();
            if (!rmode) std::cout << "\nthreshold for sparsity = " << th << "\ncorresponding r_threshold = "
                                  << info.r_thresh << std::endl;
            char tag[64];
            std::snprintf(tag, sizeof tag, "_spa%.3f%%_cor%.3f", (double)info.s_thresh, (double)info.r_thresh);
            const std::string out = base + tag + (weighted ? "_weighted.csr" : ".csr");
            std::cout << "GPU correlation + CSR time: " << secs * 1000.0 << " ms (" << info.band_pairs
                      << " pairs rechecked in FP64)" << std::endl;
            std::cout << "generating " << out << "..." << std::endl;
            if (!pgio::write_csr(out, row.data(), n, col.data(), info.nnz, weighted ? val.data() : nullptr)) {
                std::cerr << "can't write " << out << "\n";
                return 2;
            }
            pgio::time_log(std::string("CorMat\t") + a + "GPU\tkernel time\t" + std::to_string(secs) + "s");
        }
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
