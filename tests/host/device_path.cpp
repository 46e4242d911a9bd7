// device_path.cpp -- a host build of the library's device-side quality and
// optimiser code (gm_quality.h, gm_smoothers.h) for the texture-sampler
// cases, run by tests/test_device_headers_on_host.py under the address and
// undefined-behaviour sanitizers.
//
// Reads raw arrays from the directory given as argv[1] (written by the test)
// and runs, on the renumbered hexahedral mesh:
//   * two Nelder-Mead passes with the metric-conformity evaluator and the
//     texture sampler, one vertex at a time in colour order, each patch
//     evaluated by one loop over its elements (the one-lane-per-vertex
//     kernel's functor);
//   * the GETMe hexahedron element step of the pass-start positions (the
//     element kernel's arithmetic, metric-wise volumes through the adaptive
//     walk), counting the walk's iterations per element.
// Writes nm_pos.bin, nm_counts.bin, getme_vpp.bin, getme_w.bin and
// walk_iters.bin back into the same directory.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

static long long g_walkSteps = 0;
#define GMQ_WALK_STEP() (++g_walkSteps)
#include "../../gpumesh_amd/csrc/gm_quality.h"
#include "../../gpumesh_amd/csrc/gm_smoothers.h"

using namespace gmq;
typedef v3<double> V;

template<typename X> static std::vector<X> load(const std::string& dir, const char* name)
{
    std::vector<X> v;
    FILE* f = std::fopen((dir + "/" + name).c_str(), "rb");
    if(!f) { std::fprintf(stderr, "missing %s\n", name); std::exit(2); }
    X x;
    while(std::fread(&x, sizeof(X), 1, f) == 1) v.push_back(x);
    std::fclose(f);
    return v;
}

template<typename X> static void save(const std::string& dir, const char* name, const std::vector<X>& v)
{
    FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
    std::fwrite(v.data(), sizeof(X), v.size(), f);
    std::fclose(f);
}

struct HexMesh
{
    std::vector<V> p;
    std::vector<uint32_t> hexs, eoff, eref, noff, nbr;
};

// The one-lane-per-vertex kernel's patch functor on the host.
struct Patch
{
    HexMesh* m;
    const Crew* c;
    uint32_t v;
    int count;
    double operator()(V x)
    {
        ++count;
        double pq = 0.0, pw = 0.0;
        for(uint32_t r = m->eoff[v]; r < m->eoff[v + 1]; ++r)
        {
            const uint32_t ref = m->eref[r], id = ref >> 4;
            const int k = (int)(ref & 7);
            V vp[8];
            for(int j = 0; j < 8; ++j) vp[j] = (j == k) ? x : m->p[m->hexs.at(8 * (size_t)id + j)];
            accumulatePatch<double>(pq, pw, hexQuality<double, EV_METRIC_CONFORMITY, SA_TEXTURE>(*c, vp));
        }
        return finalizePatch<double>(pq, pw);
    }
};

int main(int argc, char** argv)
{
    if(argc < 2) return 2;
    const std::string dir = argv[1];
    HexMesh m;
    const std::vector<double> pos = load<double>(dir, "pos.bin");
    m.hexs = load<uint32_t>(dir, "hexs.bin");
    m.eoff = load<uint32_t>(dir, "eoff.bin");
    m.eref = load<uint32_t>(dir, "eref.bin");
    m.noff = load<uint32_t>(dir, "noff.bin");
    m.nbr = load<uint32_t>(dir, "nbr.bin");
    const std::vector<uint32_t> order = load<uint32_t>(dir, "order.bin");
    const std::vector<int32_t> tsize = load<int32_t>(dir, "tsize.bin");
    const std::vector<double> box = load<double>(dir, "box.bin");      // tmin[3], text[3], K
    const std::vector<float> texels = load<float>(dir, "texels.bin");
    const size_t V_ = pos.size() / 3;
    m.p.resize(V_);
    for(size_t i = 0; i < V_; ++i) m.p[i] = mk<double>(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);

    Crew c{};
    const double K = box.at(6);
    c.scaling = K;
    c.scaling2 = K * K;
    c.scaling3 = K * K * K;
    for(int d = 0; d < 3; ++d) { c.tsize[d] = tsize.at(d); c.tmin[d] = box[d]; c.text[d] = box[3 + d]; }
    c.texels = texels.data();

    // GETMe hexahedron element step (pass-start positions).
    const size_t H = m.hexs.size() / 8;
    std::vector<double> vppOut(24 * H), wOut(H);
    std::vector<int64_t> iters(H);
    for(size_t e = 0; e < H; ++e)
    {
        V vp[8], vpp[8];
        for(int k = 0; k < 8; ++k) vp[k] = m.p[m.hexs[8 * e + k]];
        V aux[6] = {(vp[0] + vp[1] + vp[2] + vp[3]) / 4.0, (vp[0] + vp[4] + vp[5] + vp[1]) / 4.0,
                    (vp[1] + vp[5] + vp[6] + vp[2]) / 4.0, (vp[2] + vp[6] + vp[7] + vp[3]) / 4.0,
                    (vp[0] + vp[3] + vp[7] + vp[4]) / 4.0, (vp[4] + vp[7] + vp[6] + vp[5]) / 4.0};
        V n[8] = {cross(aux[1] - aux[0], aux[4] - aux[0]), cross(aux[2] - aux[0], aux[1] - aux[0]),
                  cross(aux[3] - aux[0], aux[2] - aux[0]), cross(aux[4] - aux[0], aux[3] - aux[0]),
                  cross(aux[4] - aux[5], aux[1] - aux[5]), cross(aux[1] - aux[5], aux[2] - aux[5]),
                  cross(aux[2] - aux[5], aux[3] - aux[5]), cross(aux[3] - aux[5], aux[4] - aux[5])};
        V bases[8] = {(aux[0] + aux[1] + aux[4]) / 3.0, (aux[0] + aux[2] + aux[1]) / 3.0,
                      (aux[0] + aux[3] + aux[2]) / 3.0, (aux[0] + aux[4] + aux[3]) / 3.0,
                      (aux[5] + aux[4] + aux[1]) / 3.0, (aux[5] + aux[1] + aux[2]) / 3.0,
                      (aux[5] + aux[2] + aux[3]) / 3.0, (aux[5] + aux[3] + aux[4]) / 3.0};
        for(int k = 0; k < 8; ++k) vpp[k] = bases[k] + n[k] * tdiv(0.78, tsqrt(tsqrt(dot(n[k], n[k]))));
        g_walkSteps = 0;
        const double vol = hexVolume<double, SA_TEXTURE>(c, vp);
        const double volp = hexVolume<double, SA_TEXTURE>(c, vpp);
        iters[e] = g_walkSteps;
        const double var = tcbrt(tabs(tdiv(vol, volp)));
        V center = (1.0 / 8.0) * (vp[0] + vp[1] + vp[2] + vp[3] + vp[4] + vp[5] + vp[6] + vp[7]);
        for(int k = 0; k < 8; ++k) vpp[k] = center + var * (vpp[k] - center);
        for(int k = 0; k < 8; ++k) { vppOut[24 * e + 3 * k] = vpp[k].x; vppOut[24 * e + 3 * k + 1] = vpp[k].y; vppOut[24 * e + 3 * k + 2] = vpp[k].z; }
        wOut[e] = hexQuality<double, EV_METRIC_CONFORMITY, SA_TEXTURE>(c, vpp) /
                  (1.0 + hexQuality<double, EV_METRIC_CONFORMITY, SA_TEXTURE>(c, vp));
    }
    save(dir, "getme_vpp.bin", vppOut);
    save(dir, "getme_w.bin", wOut);
    save(dir, "walk_iters.bin", iters);

    // Two Nelder-Mead passes, colour order.
    std::vector<int32_t> counts(V_, 0);
    for(int pass = 0; pass < 2; ++pass)
        for(uint32_t v : order)
        {
            const V p0 = m.p[v];
            double total = 0.0;
            for(uint32_t r = m.noff[v]; r < m.noff[v + 1]; ++r)
            {
                const V d = p0 - m.p[m.nbr[r]];
                total += tsqrt(dot(d, d));
            }
            Patch f{&m, &c, v, 0};
            m.p[v] = gms::nelderMead<double>(p0, total / (double)(m.noff[v + 1] - m.noff[v]), f);
            counts[v] += f.count;
        }
    std::vector<double> out(3 * V_);
    for(size_t i = 0; i < V_; ++i) { out[3 * i] = m.p[i].x; out[3 * i + 1] = m.p[i].y; out[3 * i + 2] = m.p[i].z; }
    save(dir, "nm_pos.bin", out);
    save(dir, "nm_counts.bin", counts);
    return 0;
}
