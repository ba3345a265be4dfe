// CUCorMat -- voxel time series to thresholded CSR networks (drop-in for
// U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:51-378 and the weighted twin
// W/CorMat/CUCorMat/CUCorMat/CorMat.cpp).
//
//   CUCorMat [--weighted] [--fp32] Dir [Mask_Path] maskthr avg cormat {r|s} thr...
//
// Source form (mask.nii inside Dir, outputs beside the inputs) and the
// documented form with Mask_Path (outputs into Dir/unweighted or
// Dir/weighted), as SP/script.bat uses it.
//   avg   n: per-subject networks  <subject>_<N>_spa%.3f%%_cor%.3f.csr
//            (weighted: ..._weighted.csr)
//         y?: the network of the subject average only, b?: both; the second
//            character is f (average Fisher z, then map back) or n (average r)
//            (U CorMat.cpp:189, 380, 460-517).  Output <Dir>/<last dir name>_
//            <subjects>_<N>_spa%.3f%%_cor%.3f.csr (the weighted tool writes
//            .csr here too, W CorMat.cpp:587).
//   cormat y: also write the dense matrix, <subject>_<N>.cormat per subject
//            and <Dir>/<last dir name>_<subjects>_<N>.cormat for the average:
//            int32 low word of M1 = N(N-1)/2, then M1 floats, the upper triangle
//            row-major (U CorMat.cpp:300-309, 525-534).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "tool_common.hpp"

// .cormat: int32 low word of M1 (the reference writes sizeof(int) bytes of its
// long long), then the M1 floats, read from the device in 64 MiB pieces
static bool write_cormat(pgtool::Ctx &ctx, const std::string &path, int n)
{
    const long long m1 = (long long)n * (n - 1) / 2;
    std::ofstream f(path, std::ios::binary);
    const int32_t lo = (int32_t)(uint32_t)(unsigned long long)m1;
    f.write(reinterpret_cast<const char *>(&lo), 4);
    std::vector<float> buf((size_t)std::min<long long>(m1, 1ll << 24));
    for (long long off = 0; off < m1; off += (long long)buf.size()) {
        const long long cnt = std::min<long long>((long long)buf.size(), m1 - off);
        ctx.check(pg_cormat_read(ctx.c, off, cnt, buf.data()), "pg_cormat_read");
        f.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(cnt * 4));
    }
    if (!f.good()) {
        std::cerr << "can't write " << path << "\n";
        return false;
    }
    std::cout << "generating " << path << "..." << std::endl;
    return true;
}

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
                     "to_average(yf/yn/bf/bn/n) to_save_cormatrix(y/n) threshold_type(r/s) threshold(s)\n"
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
    const std::string avgs = argv[k + 1];
    const char avg = avgs[0];
    const char save = argv[k + 2][0];
    const char kind = argv[k + 3][0];
    const bool per_subject = avg == 'n' || avg == 'N' || avg == 'b' || avg == 'B';
    const bool average = avg == 'y' || avg == 'Y' || avg == 'b' || avg == 'B';
    if (!per_subject && !average) {
        std::cerr << "to_average type error!\nn, yf, yn, bf or bn\n";
        return 1;
    }
    const char ft = avgs.size() > 1 ? avgs[1] : 0;
    if (average && ft != 'f' && ft != 'F' && ft != 'n' && ft != 'N') {
        std::cerr << "to_average: the second character must be f (Fisher z) or n\n";
        return 1;
    }
    const bool fisher = ft == 'f' || ft == 'F';
    bool cormat;
    if (save == 'y' || save == 'Y') cormat = true;
    else if (save == 'n' || save == 'N') cormat = false;
    else {
        std::cerr << "to_save_cor_matrix type error! \ny to save the whole correlation matrix \nn to save only csr "
                     "format of adjacency matrix\n";
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
    std::vector<std::string> files;
    for (const auto &f : pgio::list_dir(dir, ".nii", "mask.nii"))
        if (!(documented && pgio::join(dir, f) == mask_path)) files.push_back(f);
    std::cout << files.size() << " file(s) to be processed." << std::endl;
    pgtool::Ctx ctx;
    pgtool::Timer total;
    for (const auto &fname : files) {
        if (!per_subject) break;
        const std::string a = pgio::join(dir, fname);
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
        if (cormat) {
            ctx.check(pg_cormat_begin(ctx.c, n, 0), "pg_cormat_begin");
            ctx.check(pg_cormat_add(ctx.c, bold.data(), t, PG_BOLD_TIME_MAJOR), "pg_cormat_add");
            ctx.check(pg_cormat_finish(ctx.c), "pg_cormat_finish");
            if (!write_cormat(ctx, base + ".cormat", n)) return 2;
        }
        // several thresholds: one Gram pass at the loosest, then one
        // selection per threshold (U CorMat.cpp:300-346 computes r once)
        const bool multi = thr.size() > 1;
        if (multi) {
            pg_corr_opts o;
            o.thr_kind = rmode ? PG_THR_R : PG_THR_SPARSITY;
            o.thr = rmode ? *std::min_element(thr.begin(), thr.end()) : *std::max_element(thr.begin(), thr.end());
            o.weighted = weighted ? 1 : 0;
            o.prec = prec;
            o.layout = PG_BOLD_TIME_MAJOR;
            pg_corr_info einfo;
            pgtool::Timer tm;
            ctx.check(pg_corr_emit(ctx.c, bold.data(), n, t, &o, &einfo), "pg_corr_emit");
            std::cout << "GPU correlation (one Gram pass for " << thr.size() << " thresholds): "
                      << tm.seconds() * 1000.0 << " ms, " << einfo.band_pairs << " pairs kept" << std::endl;
        }
        for (float th : thr) {
            pg_corr_opts o;
            o.thr_kind = rmode ? PG_THR_R : PG_THR_SPARSITY;
            o.thr = th;
            o.weighted = weighted ? 1 : 0;
            o.prec = prec;
            o.layout = PG_BOLD_TIME_MAJOR;
            pg_corr_info info;
            pgtool::Timer tm;
            if (multi) ctx.check(pg_corr_select(ctx.c, th, &info), "pg_corr_select");
            else ctx.check(pg_corr_build(ctx.c, bold.data(), n, t, &o, &info), "pg_corr_build");
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
    if (average && !files.empty()) {
        // U CorMat.cpp:380-620: every subject through the mask, one matrix
        int n = -1;
        pgtool::Timer tm;
        for (const auto &fname : files) {
            const std::string a = pgio::join(dir, fname);
            std::cout << "\ncalculating correlation for " << a << " ..." << std::endl;
            int ns = 0, t = 0;
            std::vector<float> bold;
            if (!pgio::read_bold(a, mask, prob_cut, ns, t, bold, err)) {
                std::cerr << err << "\n";
                return 2;
            }
            if (n < 0) {
                n = ns;
                ctx.check(pg_cormat_begin(ctx.c, n, fisher ? 1 : 0), "pg_cormat_begin");
            }
            ctx.check(pg_cormat_add(ctx.c, bold.data(), t, PG_BOLD_TIME_MAJOR), "pg_cormat_add");
        }
        ctx.check(pg_cormat_finish(ctx.c), "pg_cormat_finish");
        std::cout << "GPU correlation and averaging time: " << tm.seconds() * 1000.0 << " ms" << std::endl;
        // <Dir>/<last component of Dir>_<subjects>_<N> (U CorMat.cpp:519-524)
        const size_t sep = dir.find_last_of("/\\");
        const std::string last = sep == std::string::npos ? std::string() : dir.substr(sep + 1);
        const std::string base = pgio::join(outdir, last) + "_" + std::to_string(files.size()) + "_" + std::to_string(n);
        if (cormat && !write_cormat(ctx, base + ".cormat", n)) return 2;
        for (float th : thr) {
            pg_corr_info info;
            ctx.check(pg_cormat_threshold(ctx.c, rmode ? PG_THR_R : PG_THR_SPARSITY, th, weighted ? 1 : 0, &info),
                      "pg_cormat_threshold");
            std::vector<int32_t> row(n + 1), col((size_t)std::max<int64_t>(info.nnz, 1));
            std::vector<float> val(weighted ? col.size() : 0);
            ctx.check(pg_corr_csr(ctx.c, row.data(), col.data(), weighted ? val.data() : nullptr), "pg_corr_csr");
            if (!rmode) std::cout << "\nthreshold for sparsity = " << th << "\ncorresponding r_threshold = "
                                  << info.r_thresh << std::endl;
            char tag[64];
            std::snprintf(tag, sizeof tag, "_spa%.3f%%_cor%.3f", (double)info.s_thresh, (double)info.r_thresh);
            const std::string out = base + tag + ".csr";
            std::cout << "generating " << out << "..." << std::endl;
            if (!pgio::write_csr(out, row.data(), n, col.data(), info.nnz, weighted ? val.data() : nullptr)) {
                std::cerr << "can't write " << out << "\n";
                return 2;
            }
        }
    }
    std::cout << "total elapsed time: " << total.seconds() << " s." << std::endl;
    return 0;
}
