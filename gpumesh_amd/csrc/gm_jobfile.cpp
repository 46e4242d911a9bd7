// gm_jobfile.cpp -- a lossless binary job format for the serving loop.
//
// The reference stores meshes as JSON (Serialization/JsonSerializer.cpp:25-120,
// JsonDeserializer.cpp:38-115): positions printed with the stream's default
// 6 significant digits, and fixed vertices reduced to a constraint id, so a
// round trip moves vertices and loses the positions of fixed ones.  A job
// here carries the mesh bit for bit, with the smoothing request:
//
//   offset  size  field
//        0     4  magic "GMJ1"
//        4     4  header bytes (64, little-endian u32)
//        8     8  nverts (i64)
//       16     8  ntets (i64)
//       24     8  nhexs (i64)
//       32     4  smoother (i32, GM_NELDER_MEAD..GM_GETME)
//       36     4  passes (i32)
//       40     4  precision (i32, GM_FP64 / GM_FP32 / GM_FP32_POLISH / GM_FP64_FAST)
//       44    20  reserved (zero)
//       64        positions, 3 nverts f64 (x, y, z per vertex)
//                 fixed flags, nverts u8, zero-padded to a multiple of 8
//                 tetrahedra, 4 ntets u32
//                 hexahedra, 8 nhexs u32, zero-padded to a multiple of 8
//      end-8   8  FNV-1a 64 of every byte before it
//
// All integers and floats are little-endian (the byte order of x86-64 and
// AMD GPUs); the format is read and written with plain copies.
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/gpumesh.h"

namespace gmr
{
int fail(int code, const std::string& msg);
}

namespace
{

constexpr uint32_t HEADER = 64;

uint64_t pad8(uint64_t n) { return (n + 7) & ~uint64_t(7); }

bool sizes(int64_t nverts, int64_t ntets, int64_t nhexs, uint64_t& total)
{
    if(nverts < 0 || ntets < 0 || nhexs < 0) return false;
    const uint64_t lim = uint64_t(1) << 40;
    if((uint64_t)nverts > lim || (uint64_t)ntets > lim || (uint64_t)nhexs > lim) return false;
    total = HEADER + 24 * (uint64_t)nverts + pad8((uint64_t)nverts) + 16 * (uint64_t)ntets +
            pad8(32 * (uint64_t)nhexs) + 8;
    return true;
}

uint64_t fnv1a(const unsigned char* p, uint64_t n)
{
    uint64_t h = 1469598103934665603ull;
    for(uint64_t i = 0; i < n; ++i)
    {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

template<typename X> void put(unsigned char* b, uint64_t off, X v) { std::memcpy(b + off, &v, sizeof(X)); }
template<typename X> X get(const unsigned char* b, uint64_t off)
{
    X v;
    std::memcpy(&v, b + off, sizeof(X));
    return v;
}

// Validates a job's header, size and checksum.
int checkJob(const void* blob, int64_t bytes, int64_t& nv, int64_t& nt, int64_t& nh)
{
    using gmr::fail;
    if(!blob) return fail(GM_ERR_ARG, "null job");
    const unsigned char* b = (const unsigned char*)blob;
    if(bytes < (int64_t)HEADER + 8) return fail(GM_ERR_ARG, "job shorter than its header");
    if(std::memcmp(b, "GMJ1", 4) != 0) return fail(GM_ERR_ARG, "not a job (bad magic)");
    if(get<uint32_t>(b, 4) != HEADER) return fail(GM_ERR_ARG, "unknown job header size");
    nv = get<int64_t>(b, 8);
    nt = get<int64_t>(b, 16);
    nh = get<int64_t>(b, 24);
    uint64_t total = 0;
    if(!sizes(nv, nt, nh, total)) return fail(GM_ERR_ARG, "bad element counts in job");
    if((uint64_t)bytes != total) return fail(GM_ERR_ARG, "job size does not match its counts");
    if(fnv1a(b, total - 8) != get<uint64_t>(b, total - 8)) return fail(GM_ERR_ARG, "job checksum mismatch");
    return GM_OK;
}

} // namespace

extern "C" {

int64_t gm_job_size(int64_t nverts, int64_t ntets, int64_t nhexs)
{
    uint64_t total = 0;
    return sizes(nverts, ntets, nhexs, total) ? (int64_t)total : -1;
}

int gm_job_encode(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets, const uint32_t* tets,
                  int64_t nhexs, const uint32_t* hexs, int smoother, int passes, int precision, void* out,
                  int64_t outBytes)
{
    using gmr::fail;
    uint64_t total = 0;
    if(!sizes(nverts, ntets, nhexs, total)) return fail(GM_ERR_ARG, "bad element counts");
    if((nverts && (!pos || !fixed)) || (ntets && !tets) || (nhexs && !hexs) || !out)
        return fail(GM_ERR_ARG, "null array");
    if(outBytes < (int64_t)total) return fail(GM_ERR_ARG, "output buffer smaller than gm_job_size");
    if(smoother < 0 || smoother > 3 || passes < 0 || precision < 0 || precision > 3)
        return fail(GM_ERR_ARG, "bad smoothing request");
    unsigned char* b = (unsigned char*)out;
    std::memset(b, 0, total);
    std::memcpy(b, "GMJ1", 4);
    put<uint32_t>(b, 4, HEADER);
    put<int64_t>(b, 8, nverts);
    put<int64_t>(b, 16, ntets);
    put<int64_t>(b, 24, nhexs);
    put<int32_t>(b, 32, smoother);
    put<int32_t>(b, 36, passes);
    put<int32_t>(b, 40, precision);
    uint64_t off = HEADER;
    if(nverts) std::memcpy(b + off, pos, 24 * (size_t)nverts);
    off += 24 * (uint64_t)nverts;
    for(int64_t v = 0; v < nverts; ++v) b[off + v] = fixed[v] ? 1 : 0;
    off += pad8((uint64_t)nverts);
    if(ntets) std::memcpy(b + off, tets, 16 * (size_t)ntets);
    off += 16 * (uint64_t)ntets;
    if(nhexs) std::memcpy(b + off, hexs, 32 * (size_t)nhexs);
    off += pad8(32 * (uint64_t)nhexs);
    put<uint64_t>(b, off, fnv1a(b, off));
    return GM_OK;
}

int gm_job_info(const void* blob, int64_t bytes, int64_t* counts3, int32_t* request3)
{
    int64_t nv = 0, nt = 0, nh = 0;
    const int rc = checkJob(blob, bytes, nv, nt, nh);
    if(rc) return rc;
    const unsigned char* b = (const unsigned char*)blob;
    if(counts3)
    {
        counts3[0] = nv;
        counts3[1] = nt;
        counts3[2] = nh;
    }
    if(request3)
    {
        request3[0] = get<int32_t>(b, 32);
        request3[1] = get<int32_t>(b, 36);
        request3[2] = get<int32_t>(b, 40);
    }
    return GM_OK;
}

int gm_job_decode(const void* blob, int64_t bytes, double* pos, int32_t* fixed, uint32_t* tets, uint32_t* hexs)
{
    using gmr::fail;
    int64_t nv = 0, nt = 0, nh = 0;
    const int rc = checkJob(blob, bytes, nv, nt, nh);
    if(rc) return rc;
    if((nv && (!pos || !fixed)) || (nt && !tets) || (nh && !hexs)) return fail(GM_ERR_ARG, "null array");
    const unsigned char* b = (const unsigned char*)blob;
    uint64_t off = HEADER;
    if(nv) std::memcpy(pos, b + off, 24 * (size_t)nv);
    off += 24 * (uint64_t)nv;
    for(int64_t v = 0; v < nv; ++v) fixed[v] = b[off + v];
    off += pad8((uint64_t)nv);
    if(nt) std::memcpy(tets, b + off, 16 * (size_t)nt);
    off += 16 * (uint64_t)nt;
    if(nh) std::memcpy(hexs, b + off, 32 * (size_t)nh);
    return GM_OK;
}

} // extern "C"
