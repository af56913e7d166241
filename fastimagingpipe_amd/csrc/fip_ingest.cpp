// Measurement Set ingest from exported table columns: snapshot grouping,
// the per-snapshot worker pool, zero-padding and the preprocessor's output
// files (SVDPreprocessMain.cpp:337-546, 548-640).  The casacore table reader
// itself is not here: the columns (TIME, UVW, DATA, FLAG, WEIGHT_SPECTRUM,
// SPECTRAL_WINDOW/CHAN_FREQ) come from an external converter.
#include "fip.h"

#include "fip_context.hpp"
#include "fip_fits.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <future>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace fip {
namespace {

struct Group {
    std::vector<size_t> rows;   // table rows in row order
};

struct SnapOut {
    std::vector<float> bin, vin, vre, vim;
    size_t nsamp = 0;
    int rc = FIP_OK;
    std::string err;
};

struct Table {
    const double* uvw;
    const float* data;
    const unsigned char* flag;
    const float* weight;
    size_t nchan_data, npol;
    const double* freqs;
    size_t nchan;
};

// Snapshots in the order the reference forms them: table files in
// ascending file index; within a file one snapshot, or with TIME grouping
// one per distinct TIME in ascending order (group_rows_by_time,
// SVDPreprocessMain.cpp:387-409), each keeping its rows in row order.
std::vector<Group> group_rows(const int* file_index, const double* time, size_t nrows, bool by_time)
{
    std::map<int, std::map<double, std::vector<size_t>>> files;
    for (size_t r = 0; r < nrows; ++r) {
        const int f = file_index ? file_index[r] : 0;
        const double t = by_time ? time[r] : 0.0;
        files[f][t].push_back(r);
    }
    std::vector<Group> out;
    for (auto& f : files)
        for (auto& t : f.second) out.push_back(Group{std::move(t.second)});
    return out;
}

// One snapshot: XX (pol 0) and YY (pol 3) of every SPECTRAL_WINDOW channel c
// from DATA channel c % nchan_data, flags and weights (1 without
// WEIGHT_SPECTRUM), flattened channel-major (n = row + c * nrows,
// SVDPreprocessMain.cpp:430-536); UVW and frequencies are held as float as
// the reference's reader stores them.  Then the weighted collapse and the
// SVD-plane projection (fip_collapse, fip_project).
void process(const Table& T, const Group& g, SnapOut& o)
{
    const size_t nr = g.rows.size(), nc = T.nchan, ns = nr * nc;
    std::vector<double> uvw(nr * 3), freqs(nc);
    for (size_t k = 0; k < nr; ++k)
        for (int a = 0; a < 3; ++a) uvw[3 * k + a] = (double)(float)T.uvw[3 * g.rows[k] + a];
    for (size_t c = 0; c < nc; ++c) freqs[c] = (double)(float)T.freqs[c];
    std::vector<float> xr(ns), xi(ns), yr(ns), yi(ns), w0(ns, 1.0f), w3(ns, 1.0f);
    std::vector<unsigned char> f0(ns), f3(ns);
    for (size_t k = 0; k < nr; ++k) {
        const size_t row = g.rows[k];
        for (size_t c = 0; c < nc; ++c) {
            const size_t n = k + c * nr;
            const size_t base = (row * T.nchan_data + c % T.nchan_data) * T.npol;
            xr[n] = T.data[2 * base];
            xi[n] = T.data[2 * base + 1];
            yr[n] = T.data[2 * (base + 3)];
            yi[n] = T.data[2 * (base + 3) + 1];
            f0[n] = T.flag[base] ? 1 : 0;
            f3[n] = T.flag[base + 3] ? 1 : 0;
            if (T.weight) {
                w0[n] = T.weight[base];
                w3[n] = T.weight[base + 3];
            }
        }
    }
    o.nsamp = ns;
    o.vre.resize(ns);
    o.vim.resize(ns);
    o.bin.resize(2 * ns);
    o.vin.resize(9);
    o.rc = fip_collapse(xr.data(), xi.data(), yr.data(), yi.data(), f0.data(), f3.data(), w0.data(), w3.data(), ns,
                        o.vre.data(), o.vim.data());
    if (o.rc == FIP_OK) o.rc = fip_project(uvw.data(), nr, freqs.data(), nc, o.bin.data(), o.vin.data());
    if (o.rc != FIP_OK) o.err = fip_last_error();
}

}  // namespace
}  // namespace fip

using namespace fip;

extern "C" int fip_ingest_table(const int* file_index, const double* time, const double* uvw, size_t nrows,
                                const float* data, const unsigned char* flag, const float* weight,
                                size_t nchan_data, size_t npol, const double* freqs, size_t nchan,
                                int group_by_time, int workers, const char* prefix, const char* combined,
                                size_t* nsnap_out, size_t* nsamp_out)
{
    try {
        if (!uvw || !data || !flag || !freqs || (group_by_time && !time))
            return set_error(FIP_ERR_ARG, "null pointer argument");
        if (nrows == 0) return set_error(FIP_ERR_ARG, "no rows");
        if (nchan_data == 0 || nchan == 0) return set_error(FIP_ERR_ARG, "DATA and CHAN_FREQ need channels");
        if (npol < 4) return set_error(FIP_ERR_ARG, "DATA has fewer than 4 polarisations");
        if (workers < 0) return set_error(FIP_ERR_ARG, "workers must be >= 0");
        if (group_by_time)
            for (size_t r = 0; r < nrows; ++r)
                if (!std::isfinite(time[r])) return set_error(FIP_ERR_ARG, "non-finite TIME in row " + std::to_string(r));
        const Table T{uvw, data, flag, weight, nchan_data, npol, freqs, nchan};
        const std::vector<Group> groups = group_rows(file_index, time, nrows, group_by_time != 0);
        const size_t S = groups.size();
        // resolve_snapshot_workers (SVDPreprocessMain.cpp:337-349)
        size_t nw = (size_t)workers;
        if (nw == 0) {
            const unsigned hw = std::thread::hardware_concurrency();
            nw = std::min<size_t>(4, std::min<size_t>(hw == 0 ? 2 : hw, S));
        }
        nw = std::max<size_t>(1, std::min(nw, S));
        std::vector<SnapOut> out(S);
        std::deque<std::future<void>> running;
        size_t next = 0;
        while (next < S || !running.empty()) {
            while (next < S && running.size() < nw) {
                const size_t k = next++;
                running.push_back(std::async(std::launch::async, [&T, &groups, &out, k]() { process(T, groups[k], out[k]); }));
            }
            running.front().get();
            running.pop_front();
        }
        for (size_t k = 0; k < S; ++k)
            if (out[k].rc != FIP_OK)
                return set_error(out[k].rc, "snapshot " + std::to_string(k) + ": " + out[k].err);
        // Zero-pad every snapshot to the largest sample count
        // (append_padded_outputs, SVDPreprocessMain.cpp:264-291).
        size_t np = 0;
        for (const SnapOut& o : out) np = std::max(np, o.nsamp);
        std::vector<float> bin(S * np * 2, 0.0f), vin(S * 9), vre(S * np, 0.0f), vim(S * np, 0.0f), vis(S * np * 2, 0.0f);
        for (size_t k = 0; k < S; ++k) {
            const SnapOut& o = out[k];
            std::copy(o.bin.begin(), o.bin.end(), bin.begin() + k * np * 2);
            std::copy(o.vin.begin(), o.vin.end(), vin.begin() + k * 9);
            std::copy(o.vre.begin(), o.vre.end(), vre.begin() + k * np);
            std::copy(o.vim.begin(), o.vim.end(), vim.begin() + k * np);
            for (size_t n = 0; n < o.nsamp; ++n) {
                vis[2 * (k * np + n)] = o.vre[n];
                vis[2 * (k * np + n) + 1] = o.vim[n];
            }
        }
        // The four 2-D files and the combined 3-extension file
        // (SVDPreprocessMain.cpp:610-624, 181-240).
        const std::string pre = prefix ? prefix : "";
        const long tot = (long)(S * np);
        if (int rc = fip_fits_write((pre + "Bin.fits").c_str(), bin.data(), 2, tot)) return rc;
        if (int rc = fip_fits_write((pre + "Vin.fits").c_str(), vin.data(), 3, (long)(vin.size() / 3))) return rc;
        if (int rc = fip_fits_write((pre + "Visreal.fits").c_str(), vre.data(), 1, tot)) return rc;
        if (int rc = fip_fits_write((pre + "Visimag.fits").c_str(), vim.data(), 1, tot)) return rc;
        const std::string comb = (combined && *combined) ? std::string(combined) : pre + "input.fits";
        const FitsImage ext[3] = {{"TRANSFORMATION MATRICES", 3, {3, 3, (long)S}, vin.data()},
                                  {"VISIBILITIES", 3, {2, (long)np, (long)S}, vis.data()},
                                  {"R-COORDINATES", 3, {2, (long)np, (long)S}, bin.data()}};
        if (int rc = fits_write_extensions(comb.c_str(), ext, 3)) return rc;
        if (nsnap_out) *nsnap_out = S;
        if (nsamp_out) *nsamp_out = np;
        return FIP_OK;
    } catch (const std::bad_alloc&) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}
