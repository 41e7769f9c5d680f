// Hierarchy persistence: the fingerprint of what a build depends on, and the hierarchy's structure
// (its levels; no matrix values) on disk. Replaces amgHierarchySignature (device_amg.cu:1262-1285) and
// writeAMGCache/readAMGCache (device_amg_cache.cu:176-326) with a format of this library's own.
//
// File: "MFVAMG01", u64 signature, i32 kind, target, mergeLevels, nLevels, then per level i32
// nFine, nFineFaces, nCoarse, nCoarseFaces and the arrays map[nFine], cOwn[nCoarseFaces],
// cNei[nCoarseFaces], faceRestrict[nFineFaces], i8 faceFlip[nFineFaces]; last a u64 checksum of every
// byte before it. Little-endian, as the host is. A write goes to "<path>.tmp.<pid>" and is renamed
// over the path, so a reader sees the old file or the new one, never half of one.
#include "amg.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <unistd.h>

namespace mfv {

namespace {

constexpr char kMagic[8] = {'M', 'F', 'V', 'A', 'M', 'G', '0', '1'};
constexpr uint64_t kBuildVersion = 1;   // bump when a build's result changes for the same inputs

// 64-bit fingerprint: every 8-byte word mixed in with a multiply-xorshift step, each array closed by
// its length, so arrays of different lengths cannot collide by concatenation
struct Fingerprint
{
    uint64_t h = 0x6a09e667f3bcc909ull;
    void word(uint64_t w)
    {
        h ^= w + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0xbf58476d1ce4e5b9ull;
        h ^= h >> 31;
    }
    void bytes(const void* p, size_t n)
    {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= n; i += 8)
        {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            word(w);
        }
        if (i < n)
        {
            uint64_t w = 0;
            std::memcpy(&w, b + i, n - i);
            word(w);
        }
        word((uint64_t)n);
    }
    template <typename T>
    void value(T v) { bytes(&v, sizeof(T)); }
    template <typename T>
    void list(const std::vector<T>& v) { bytes(v.data(), v.size() * sizeof(T)); }
};

struct Writer
{
    std::string buf;
    template <typename T>
    void value(T v) { buf.append(reinterpret_cast<const char*>(&v), sizeof(T)); }
    template <typename T>
    void list(const std::vector<T>& v) { buf.append(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T)); }
};

struct Reader
{
    const std::string& buf;
    size_t at = 0;
    bool ok = true;
    template <typename T>
    T value()
    {
        T v{};
        if (at + sizeof(T) > buf.size()) { ok = false; return v; }
        std::memcpy(&v, buf.data() + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    template <typename T>
    void list(std::vector<T>& v, long long n)
    {
        if (n < 0 || at + (size_t)n * sizeof(T) > buf.size()) { ok = false; return; }
        v.resize((size_t)n);
        if (n) std::memcpy(v.data(), buf.data() + at, (size_t)n * sizeof(T));
        at += (size_t)n * sizeof(T);
    }
};

uint64_t checksum(const char* p, size_t n)
{
    Fingerprint f;
    f.bytes(p, n);
    return f.h;
}

} // namespace

uint64_t amgSignature(const std::vector<int>& owner, const std::vector<int>& neighbour,
                      const std::vector<double>& weights, int nCells, const AmgBuildControls& ctl)
{
    Fingerprint f;
    f.value(kBuildVersion);
    f.value((uint64_t)sizeof(int));
    f.value((uint64_t)sizeof(double));
    f.value((int64_t)nCells);
    f.value((int64_t)ctl.kind);
    f.value((int64_t)ctl.target);
    f.value((int64_t)ctl.mergeLevels);
    f.list(owner);
    f.list(neighbour);
    f.list(weights);
    return f.h;
}

void amgSave(const Amg& a, uint64_t signature, const std::string& path)
{
    Writer w;
    w.buf.append(kMagic, sizeof(kMagic));
    w.value(signature);
    w.value((int32_t)a.controls.kind);
    w.value((int32_t)a.controls.target);
    w.value((int32_t)a.controls.mergeLevels);
    w.value((int32_t)a.agg.size());
    for (const AggLevel& L : a.agg)
    {
        w.value((int32_t)L.nFine);
        w.value((int32_t)L.nFineFaces);
        w.value((int32_t)L.nCoarse);
        w.value((int32_t)L.nCoarseFaces);
        w.list(L.map);
        w.list(L.cOwn);
        w.list(L.cNei);
        w.list(L.faceRestrict);
        w.list(L.faceFlip);
    }
    w.value(checksum(w.buf.data(), w.buf.size()));
    const std::string tmp = path + ".tmp." + std::to_string((long long)getpid());
    {
        std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
        if (!out) throw Error(kErrInvalid, "mfv: amg save: cannot write " + tmp);
        out.write(w.buf.data(), (std::streamsize)w.buf.size());
        out.flush();
        if (!out) throw Error(kErrInvalid, "mfv: amg save: write failed on " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0)
    {
        std::remove(tmp.c_str());
        throw Error(kErrInvalid, "mfv: amg save: cannot rename onto " + path);
    }
}

CacheRead amgReadLevels(const std::string& path, uint64_t signature, std::vector<AggLevel>& levels,
                        AmgBuildControls& ctl)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) return CacheRead::Absent;
    const std::string buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (buf.size() < sizeof(kMagic) + 16 || std::memcmp(buf.data(), kMagic, sizeof(kMagic)) != 0)
        return CacheRead::Unreadable;
    uint64_t stored;
    std::memcpy(&stored, buf.data() + buf.size() - 8, 8);
    if (stored != checksum(buf.data(), buf.size() - 8)) return CacheRead::Unreadable;
    Reader r{buf, sizeof(kMagic)};
    const uint64_t sig = r.value<uint64_t>();
    if (sig != signature) return CacheRead::OtherMeshOrBuild;
    AmgBuildControls c;
    c.kind = (AggKind)r.value<int32_t>();
    c.target = r.value<int32_t>();
    c.mergeLevels = r.value<int32_t>();
    const int32_t nL = r.value<int32_t>();
    if (!r.ok || nL < 0 || nL > 64) return CacheRead::Unreadable;
    std::vector<AggLevel> out((size_t)nL);
    for (AggLevel& L : out)
    {
        L.nFine = r.value<int32_t>();
        L.nFineFaces = r.value<int32_t>();
        L.nCoarse = r.value<int32_t>();
        L.nCoarseFaces = r.value<int32_t>();
        if (!r.ok) return CacheRead::Unreadable;
        r.list(L.map, L.nFine);
        r.list(L.cOwn, L.nCoarseFaces);
        r.list(L.cNei, L.nCoarseFaces);
        r.list(L.faceRestrict, L.nFineFaces);
        r.list(L.faceFlip, L.nFineFaces);
        if (!r.ok) return CacheRead::Unreadable;
    }
    if (r.at + 8 != buf.size()) return CacheRead::Unreadable;
    levels = std::move(out);
    ctl = c;
    return CacheRead::Loaded;
}

} // namespace mfv
