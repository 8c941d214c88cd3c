// ddfan_dicom.cpp -- DICOM-CT-PD projection files and CT image headers.
//
// What the reference reads with MATLAB's dicominfo/dicomread and its cov_*
// byte converters (readProj.m:31-56, CollectReconCfg.m:5-92,
// CollectImageCfg.m:3-31, cov_4uint8_to_float.cpp, cov_8uint8_to_double.cpp,
// cov_8uint8_to_2floats.cpp, cov_uint8_to_double.cpp,
// cov_uint8_to_PhoStat.cpp): a little-endian DICOM parser (explicit or
// implicit VR; sequences skipped, nested or of undefined length), decoders of
// the top-level elements the reconstruction needs, and a threaded reader of
// the 16-bit projection frames.  Private elements are matched by their tag
// number, as the reference's dictionary does, and decoded from their bytes
// the way its converters do (little-endian float32 / float64 / uint16, and
// decimal strings through strtod).
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ddfan.h"
#include "ddfan_guard.h"

namespace ddf {
int fail(int code, const char* fmt, ...);
}

namespace {

using ddf::fail;

constexpr uint32_t tag(uint16_t g, uint16_t e) { return (uint32_t)g << 16 | e; }
constexpr uint32_t kPixelData = tag(0x7FE0, 0x0010);

struct Elem {
    char vr[3] = {0, 0, 0};
    size_t off = 0;  // value offset in the file
    uint32_t len = 0;
    bool undefined = false;  // sequence or encapsulated data of undefined length
};

struct Dicom {
    std::vector<uint8_t> buf;
    std::string path, err;
    bool explicitVR = true, native = true;
    std::map<uint32_t, Elem> top;

    uint16_t u16(size_t o) const { return (uint16_t)(buf[o] | buf[o + 1] << 8); }
    uint32_t u32(size_t o) const {
        return (uint32_t)buf[o] | (uint32_t)buf[o + 1] << 8 | (uint32_t)buf[o + 2] << 16 |
               (uint32_t)buf[o + 3] << 24;
    }
    bool bad(const char* what, size_t o) {
        char m[160];
        snprintf(m, sizeof m, "%s at byte %zu", what, o);
        err = m;
        return false;
    }

    static bool long_vr(const char* vr) {
        static const char* kLong[] = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ",
                                      "SV", "UC", "UN", "UR", "UT", "UV"};
        for (const char* l : kLong)
            if (vr[0] == l[0] && vr[1] == l[1]) return true;
        return false;
    }

    // One element header at o: tag, VR (explicit), value length and offset.
    bool header(size_t& o, bool expl, uint32_t& t, Elem& e) {
        if (o + 8 > buf.size()) return bad("truncated element header", o);
        t = tag(u16(o), u16(o + 2));
        if ((t >> 16) == 0xFFFE || !expl) {  // items and delimiters carry no VR
            e.len = u32(o + 4);
            o += 8;
        } else {
            e.vr[0] = (char)buf[o + 4];
            e.vr[1] = (char)buf[o + 5];
            if (long_vr(e.vr)) {
                if (o + 12 > buf.size()) return bad("truncated element header", o);
                e.len = u32(o + 8);
                o += 12;
            } else {
                e.len = u16(o + 6);
                o += 8;
            }
        }
        e.off = o;
        e.undefined = e.len == 0xFFFFFFFFu;
        return true;
    }

    // Skips the items of a sequence whose value starts at o (defined length
    // `len`, or up to its delimitation item); nested sequences recurse.
    bool skip_sequence(size_t& o, bool undefinedLen, uint32_t len, bool expl, int depth) {
        if (depth > 16) return bad("sequences nested too deeply", o);
        const size_t end = undefinedLen ? buf.size() : o + len;
        if (end > buf.size()) return bad("sequence runs past the end of the file", o);
        while (o < end) {
            uint32_t t;
            Elem it;
            if (!header(o, expl, t, it)) return false;
            if (t == tag(0xFFFE, 0xE0DD)) return true;  // sequence delimitation
            if (t != tag(0xFFFE, 0xE000)) return bad("expected a sequence item", o - 8);
            if (!it.undefined) {
                if (o + it.len > buf.size()) return bad("item runs past the end of the file", o);
                o += it.len;
                continue;
            }
            for (;;) {  // item of undefined length: elements up to its delimitation
                uint32_t et;
                Elem el;
                if (!header(o, expl, et, el)) return false;
                if (et == tag(0xFFFE, 0xE00D)) break;
                if (!value(o, et, el, expl, depth)) return false;
            }
        }
        if (!undefinedLen && o != end) return bad("sequence length mismatch", o);
        return undefinedLen ? bad("sequence without delimitation", o) : true;
    }

    // Steps over the value of element (t, e) whose value starts at o.  A value
    // of defined length (sequences included) is skipped whole; one of undefined
    // length is a sequence (an undefined-length UN one in implicit VR) or, for
    // the pixel data, encapsulated frames.
    bool value(size_t& o, uint32_t t, const Elem& e, bool expl, int depth) {
        if (e.undefined) {
            if (t == kPixelData) {  // encapsulated (compressed) frames
                native = false;
                o = buf.size();
                return true;
            }
            const bool un = e.vr[0] == 'U' && e.vr[1] == 'N';
            return skip_sequence(o, true, 0, expl && !un, depth + 1);
        }
        if (o + e.len > buf.size()) return bad("value runs past the end of the file", o);
        o += e.len;
        return true;
    }

    // Reads and indexes the file.  stopAtPixels: the header only.
    bool load(const char* p, bool stopAtPixels) {
        path = p;
        FILE* f = fopen(p, "rb");
        if (!f) return bad("cannot open file", 0);
        if (fseek(f, 0, SEEK_END) == 0) {
            const long n = ftell(f);
            if (n > 0) {
                buf.resize((size_t)n);
                rewind(f);
                if (fread(buf.data(), 1, buf.size(), f) != buf.size()) buf.clear();
            }
        }
        fclose(f);
        if (buf.size() < 132 || memcmp(&buf[128], "DICM", 4) != 0)
            return bad("not a DICOM file (no DICM prefix)", 128);
        size_t o = 132;
        std::string ts;
        while (o + 8 <= buf.size() && u16(o) == 0x0002) {  // file meta: explicit VR LE
            uint32_t t;
            Elem e;
            if (!header(o, true, t, e)) return false;
            if (e.undefined || o + e.len > buf.size()) return bad("bad file meta element", o);
            if (t == tag(0x0002, 0x0010)) ts.assign((const char*)&buf[o], e.len);
            o += e.len;
        }
        while (!ts.empty() && (ts.back() == '\0' || ts.back() == ' ')) ts.pop_back();
        if (ts.empty()) return bad("no transfer syntax in the file meta information", o);
        if (ts == "1.2.840.10008.1.2") {
            explicitVR = false;
        } else if (ts == "1.2.840.10008.1.2.2" || ts == "1.2.840.10008.1.2.1.99") {
            return bad(("unsupported transfer syntax " + ts).c_str(), o);
        } else if (ts != "1.2.840.10008.1.2.1") {
            native = false;  // compressed pixel data; the data set is explicit VR LE
        }
        while (o < buf.size()) {
            uint32_t t;
            Elem e;
            if (!header(o, explicitVR, t, e)) return false;
            if ((t >> 16) == 0xFFFE) return bad("unexpected item at the top level", o - 8);
            top[t] = e;
            if (t == kPixelData && stopAtPixels) return true;
            if (!value(o, t, e, explicitVR, 0)) return false;
        }
        return true;
    }

    const Elem* find(uint32_t t) const {
        auto it = top.find(t);
        return it == top.end() || it->second.undefined ? nullptr : &it->second;
    }
    bool f32(uint32_t t, int idx, float* v) const {
        const Elem* e = find(t);
        if (!e || e->len < 4u * (idx + 1)) return false;
        const uint32_t b = u32(e->off + 4 * idx);
        memcpy(v, &b, 4);
        return true;
    }
    bool f64(uint32_t t, int idx, double* v) const {
        const Elem* e = find(t);
        if (!e || e->len < 8u * (idx + 1)) return false;
        const uint64_t b = (uint64_t)u32(e->off + 8 * idx) |
                           (uint64_t)u32(e->off + 8 * idx + 4) << 32;
        memcpy(v, &b, 8);
        return true;
    }
    bool us(uint32_t t, int* v) const {
        const Elem* e = find(t);
        if (!e || e->len < 2) return false;
        *v = u16(e->off);
        return true;
    }
    std::string str(uint32_t t) const {
        const Elem* e = find(t);
        if (!e) return std::string();
        std::string s((const char*)&buf[e->off], e->len);
        while (!s.empty() && (s.back() == '\0' || s.back() == ' ')) s.pop_back();
        size_t b = 0;
        while (b < s.size() && s[b] == ' ') ++b;
        return s.substr(b);
    }
    // Value idx of a decimal / integer string (backslash-separated).
    bool ds(uint32_t t, int idx, double* v) const {
        const std::string s = str(t);
        if (s.empty()) return false;
        size_t b = 0;
        for (int k = 0; k < idx; ++k) {
            b = s.find('\\', b);
            if (b == std::string::npos) return false;
            ++b;
        }
        const std::string item = s.substr(b, s.find('\\', b) - b);
        char* end = nullptr;
        const double x = strtod(item.c_str(), &end);
        if (end == item.c_str()) return false;
        *v = x;
        return true;
    }
    int count(uint32_t t, int unit) const {
        const Elem* e = find(t);
        return e ? (int)(e->len / unit) : 0;
    }
};

int io_fail(const Dicom& d) {
    return fail(DDFAN_EIO, "%s: %s", d.path.c_str(), d.err.c_str());
}

int missing(const Dicom& d, uint32_t t) {
    return fail(DDFAN_EIO, "%s: missing or short element (%04X,%04X)", d.path.c_str(), t >> 16,
                t & 0xFFFF);
}

// Pixel samples of one file into out (rows*cols 16-bit values).
bool read_frame(Dicom& d, const char* path, int rows, int cols, uint16_t* out) {
    d = Dicom();
    if (!d.load(path, false)) return false;
    int r = 0, c = 0, bits = 0, spp = 1;
    if (!d.us(tag(0x0028, 0x0010), &r) || !d.us(tag(0x0028, 0x0011), &c) ||
        !d.us(tag(0x0028, 0x0100), &bits))
        return d.bad("missing Rows, Columns or BitsAllocated", 0);
    d.us(tag(0x0028, 0x0002), &spp);
    double frames = 1.0;
    d.ds(tag(0x0028, 0x0008), 0, &frames);
    if (r != rows || c != cols) {
        char m[128];
        snprintf(m, sizeof m, "frame is %d x %d, expected %d x %d", r, c, rows, cols);
        return d.bad(m, 0);
    }
    if (bits != 16 || spp != 1 || frames != 1.0)
        return d.bad("pixel data must be one frame of 16-bit single samples", 0);
    if (!d.native) return d.bad("compressed pixel data is not supported", 0);
    const Elem* px = d.find(kPixelData);
    const size_t n = (size_t)rows * cols;
    if (!px || px->len < 2 * n) return d.bad("pixel data missing or short", 0);
    for (size_t k = 0; k < n; ++k) out[k] = d.u16(px->off + 2 * k);
    return true;
}

}  // namespace

extern "C" {

int ddfan_ctpd_read_header(const char* path, ddfan_ctpd_header* out, float* photon_stats,
                           int capacity) {
    return ddf::guarded([&]() -> int {
        if (!path || !out || capacity < 0 || (capacity > 0 && !photon_stats))
            return fail(DDFAN_EINVAL, "ddfan_ctpd_read_header: bad arguments");
        Dicom d;
        if (!d.load(path, true)) return io_fail(d);
        ddfan_ctpd_header h;
        memset(&h, 0, sizeof h);
        const double nan = std::nan("");
        // required geometry (CollectReconCfg.m:22-79)
        struct F32 {
            uint32_t t;
            int idx;
            float* dst;
        } req[] = {
            {tag(0x7029, 0x1002), 0, &h.DetectorElementTransverseSpacing},
            {tag(0x7029, 0x1006), 0, &h.DetectorElementAxialSpacing},
            {tag(0x7031, 0x1001), 0, &h.DetectorFocalCenterAngularPosition},
            {tag(0x7031, 0x1003), 0, &h.DetectorFocalCenterRadialDistance},
            {tag(0x7031, 0x1033), 0, &h.DetectorCentralElementX},
            {tag(0x7031, 0x1033), 1, &h.DetectorCentralElementY},
            {tag(0x7031, 0x1031), 0, &h.ConstantRadialDistance},
        };
        for (const F32& r : req)
            if (!d.f32(r.t, r.idx, r.dst)) return missing(d, r.t);
        if (!d.us(tag(0x7029, 0x1010), &h.NumberofDetectorRows))
            return missing(d, tag(0x7029, 0x1010));
        if (!d.us(tag(0x7029, 0x1011), &h.NumberofDetectorColumns))
            return missing(d, tag(0x7029, 0x1011));
        if (!d.us(tag(0x7033, 0x1013), &h.NumberofSourceAngularSteps))
            return missing(d, tag(0x7033, 0x1013));
        if (!d.f64(tag(0x0018, 0x9311), 0, &h.SpiralPitchFactor))
            return missing(d, tag(0x0018, 0x9311));
        // optional fields: NaN / empty when absent
        F32 opt[] = {
            {tag(0x7031, 0x1002), 0, &h.DetectorFocalCenterAxialPosition},
            {tag(0x7033, 0x100B), 0, &h.SourceAngularPositionShift},
            {tag(0x7033, 0x100C), 0, &h.SourceAxialPositionShift},
            {tag(0x7033, 0x100D), 0, &h.SourceRadialDistanceShift},
        };
        for (const F32& r : opt)
            if (!d.f32(r.t, r.idx, r.dst)) *r.dst = (float)nan;
        if (!d.ds(tag(0x0018, 0x0061), 0, &h.HUCalibrationFactor)) h.HUCalibrationFactor = nan;
        if (!d.ds(tag(0x0018, 0x0090), 0, &h.DataCollectionDiameter)) h.DataCollectionDiameter = nan;
        const std::string shape = d.str(tag(0x7029, 0x100B));
        snprintf(h.DetectorShape, sizeof h.DetectorShape, "%s", shape.c_str());
        d.us(tag(0x0028, 0x0010), &h.Rows);
        d.us(tag(0x0028, 0x0011), &h.Columns);
        h.PhotonStatisticsCount = d.count(tag(0x7033, 0x1065), 4);
        for (int k = 0; k < std::min(capacity, h.PhotonStatisticsCount); ++k)
            d.f32(tag(0x7033, 0x1065), k, &photon_stats[k]);
        *out = h;
        return DDFAN_OK;
    });
}

int ddfan_ct_image_read_header(const char* path, ddfan_ct_image_header* out) {
    return ddf::guarded([&]() -> int {
        if (!path || !out) return fail(DDFAN_EINVAL, "ddfan_ct_image_read_header: bad arguments");
        Dicom d;
        if (!d.load(path, true)) return io_fail(d);
        ddfan_ct_image_header h;
        memset(&h, 0, sizeof h);
        const double nan = std::nan("");
        if (!d.us(tag(0x0028, 0x0011), &h.Width)) return missing(d, tag(0x0028, 0x0011));
        if (!d.us(tag(0x0028, 0x0010), &h.Height)) return missing(d, tag(0x0028, 0x0010));
        if (!d.ds(tag(0x0018, 0x1100), 0, &h.ReconstructionDiameter))
            return missing(d, tag(0x0018, 0x1100));
        if (!d.us(tag(0x0028, 0x0101), &h.BitDepth)) h.BitDepth = 0;
        struct DS {
            uint32_t t;
            int idx;
            double* dst;
        } dss[] = {
            {tag(0x0018, 0x0050), 0, &h.SliceThickness},
            {tag(0x0018, 0x0060), 0, &h.KVP},
            {tag(0x0018, 0x0090), 0, &h.DataCollectionDiameter},
            {tag(0x0028, 0x0030), 0, &h.PixelSpacing[0]},
            {tag(0x0028, 0x0030), 1, &h.PixelSpacing[1]},
            {tag(0x0028, 0x1050), 0, &h.WindowCenter[0]},
            {tag(0x0028, 0x1050), 1, &h.WindowCenter[1]},
            {tag(0x0028, 0x1051), 0, &h.WindowWidth[0]},
            {tag(0x0028, 0x1051), 1, &h.WindowWidth[1]},
            {tag(0x0028, 0x1052), 0, &h.RescaleIntercept},
            {tag(0x0028, 0x1053), 0, &h.RescaleSlope},
        };
        for (const DS& r : dss)
            if (!d.ds(r.t, r.idx, r.dst)) *r.dst = nan;
        struct FD {
            uint32_t t;
            int idx;
            double* dst;
        } fds[] = {
            {tag(0x0018, 0x9306), 0, &h.SingleCollimationWidth},
            {tag(0x0018, 0x9307), 0, &h.TotalCollimationWidth},
            {tag(0x0018, 0x9309), 0, &h.TableSpeed},
            {tag(0x0018, 0x9310), 0, &h.TableFeedPerRotation},
            {tag(0x0018, 0x9311), 0, &h.SpiralPitchFactor},
            {tag(0x0018, 0x9313), 0, &h.DataCollectionCenterPatient[0]},
            {tag(0x0018, 0x9313), 1, &h.DataCollectionCenterPatient[1]},
            {tag(0x0018, 0x9313), 2, &h.DataCollectionCenterPatient[2]},
            {tag(0x0018, 0x9318), 0, &h.ReconstructionTargetCenterPatient[0]},
            {tag(0x0018, 0x9318), 1, &h.ReconstructionTargetCenterPatient[1]},
            {tag(0x0018, 0x9318), 2, &h.ReconstructionTargetCenterPatient[2]},
        };
        for (const FD& r : fds)
            if (!d.f64(r.t, r.idx, r.dst)) *r.dst = nan;
        // Smallest/LargestImagePixelValue: US or SS after PixelRepresentation
        int rep = 0, lo = 0, hi = 0;
        d.us(tag(0x0028, 0x0103), &rep);
        const bool haveLo = d.us(tag(0x0028, 0x0106), &lo), haveHi = d.us(tag(0x0028, 0x0107), &hi);
        h.SmallestImagePixelValue = !haveLo ? 0 : rep ? (int)(int16_t)lo : lo;
        h.LargestImagePixelValue = !haveHi ? 0 : rep ? (int)(int16_t)hi : hi;
        h.PixelSize = h.ReconstructionDiameter / h.Width;  // CollectImageCfg.m:30
        *out = h;
        return DDFAN_OK;
    });
}

int ddfan_dicom_read_u16(const char* path, int rows, int cols, uint16_t* out) {
    return ddf::guarded([&]() -> int {
        if (!path || !out || rows < 1 || cols < 1)
            return fail(DDFAN_EINVAL, "ddfan_dicom_read_u16: bad arguments");
        Dicom d;
        if (!read_frame(d, path, rows, cols, out)) return io_fail(d);
        return DDFAN_OK;
    });
}

int ddfan_ctpd_read_frames(const char* const* paths, int nfiles, int rows, int cols, uint16_t* out,
                           int threads) {
    return ddf::guarded([&]() -> int {
        if (!paths || !out || nfiles < 0 || rows < 1 || cols < 1)
            return fail(DDFAN_EINVAL, "ddfan_ctpd_read_frames: bad arguments");
        if (nfiles == 0) return DDFAN_OK;
        for (int i = 0; i < nfiles; ++i)
            if (!paths[i]) return fail(DDFAN_EINVAL, "ddfan_ctpd_read_frames: path %d is NULL", i);
        int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
        nt = std::max(1, std::min(nt, std::min(nfiles, 64)));
        const size_t n = (size_t)rows * cols;
        std::atomic<int> next{0};
        std::mutex mu;
        int firstBad = nfiles;
        std::string firstErr;
        auto work = [&]() {
            Dicom d;
            for (int i = next++; i < nfiles; i = next++) {
                if (!read_frame(d, paths[i], rows, cols, out + (size_t)i * n)) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (i < firstBad) {
                        firstBad = i;
                        firstErr = d.path + ": " + d.err;
                    }
                    next = nfiles;  // stop handing out files
                    return;
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
        if (firstBad < nfiles) return fail(DDFAN_EIO, "%s", firstErr.c_str());
        return DDFAN_OK;
    });
}

}  // extern "C"
