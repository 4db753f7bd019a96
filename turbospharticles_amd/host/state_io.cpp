// state_io.cpp -- the reference's simulation state file (TSPSIM2) read and written from a
// tsp_state_view, so that states move between this backend and the reference's checkpoints.
//
// Layout (eval/serial/state_format.h:10-24, serialization.cpp:164-210, 860-935): a packed 64-byte
// header {magic "TSPSIM2\0", endian tag 0x01020304, header bytes 64, step u64, time f64, N_F, N_B,
// N_R, N_RB as u64}, then each population's arrays one after the other, little-endian:
//   fluid      id u64[N_F], then 16 f32 arrays in the order of TSP_F_* (VisitFluidStateFields)
//   boundary   id u64[N_B], then x, y, z, V, psi (VisitBoundaryStateFields)
//   rigid      id u64[N_R], body u32[N_R], then x, y, z offsets, V, psi (VisitRigidStateFields)
//   bodies     id u32[N_RB], particle count u32[N_RB], then the 32 f32 fields of TSP_RB_*
//              (VisitRigidBodyStateFields)
// The reader checks the magic, the endian tag, the header size, step/time and the file length
// against the header, as ReadSimStateHeader does (serialization.cpp:214-244); reader and writer
// both check the ids and the rigid memberships as SaveUniversalSimState / LoadUniversalSimState do.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "tsp_scene.h"

namespace
{
#pragma pack(push, 1)
    struct Header
    {
        char magic[8];
        uint32_t endianTag;
        uint32_t headerBytes;
        uint64_t step;
        double time;
        uint64_t nF, nB, nR, nRB;
    };
#pragma pack(pop)
    static_assert(sizeof(Header) == 64, "TSPSIM2 header");

    constexpr char kMagic[8] = { 'T', 'S', 'P', 'S', 'I', 'M', '2', '\0' };
    constexpr uint32_t kEndian = 0x01020304u;

    uint64_t fileBytes(const Header& h)
    {
        return sizeof(Header) + h.nF * (8ull + 16ull * 4ull) + h.nB * (8ull + 5ull * 4ull)
             + h.nR * (8ull + 4ull + 5ull * 4ull) + h.nRB * (2ull * 4ull + 32ull * 4ull);
    }

    struct File
    {
        std::FILE* f = nullptr;
        ~File() { if (f) { std::fclose(f); } }
    };

    int readHeader(std::FILE* f, Header& h)
    {
        if (std::fread(&h, sizeof(h), 1, f) != 1) { return TSP_ERR_INVALID; }
        if (std::memcmp(h.magic, kMagic, 8) != 0 || h.endianTag != kEndian || h.headerBytes != sizeof(Header))
        {
            return TSP_ERR_INVALID;
        }
        if (!std::isfinite(h.time) || h.time < 0.0 || h.step == std::numeric_limits<uint64_t>::max()) { return TSP_ERR_INVALID; }
        if (h.nF > 0x7fffffffull || h.nB > 0x7fffffffull || h.nR > 0x7fffffffull || h.nRB > 0xffffffffull)
        {
            return TSP_ERR_CAPACITY;
        }
        if (std::fseek(f, 0, SEEK_END) != 0) { return TSP_ERR_INVALID; }
        const long size = std::ftell(f);
        if (size < 0 || static_cast<uint64_t>(size) != fileBytes(h)) { return TSP_ERR_INVALID; }
        if (std::fseek(f, sizeof(Header), SEEK_SET) != 0) { return TSP_ERR_INVALID; }
        return TSP_OK;
    }

    // the reference's checks on a state's contents (ValidateParticleIds, serialization.cpp:245-253;
    // ValidateRigidAttachments, sync/state_sim.h:679-712): no id equals the invalid-id sentinel,
    // every rigid particle names an existing body, and each body's particle count matches its
    // members. Arrays that are null are not checked.
    bool validContents(const tsp_state_view& v)
    {
        constexpr uint64_t kInvalidId = std::numeric_limits<uint64_t>::max();
        const auto idsOk = [&](const uint64_t* ids, uint32_t n) {
            if (!ids) { return true; }
            for (uint32_t i = 0; i < n; i++) { if (ids[i] == kInvalidId) { return false; } }
            return true;
        };
        if (!idsOk(v.fluid_id, v.n_fluid) || !idsOk(v.boundary_id, v.n_boundary) || !idsOk(v.rigid_id, v.n_rigid))
        {
            return false;
        }
        if (!v.rigid_body) { return true; }
        std::vector<uint32_t> members(v.n_bodies, 0u);
        for (uint32_t i = 0; i < v.n_rigid; i++)
        {
            if (v.rigid_body[i] >= v.n_bodies) { return false; }
            members[v.rigid_body[i]]++;
        }
        if (!v.body_count) { return true; }
        for (uint32_t b = 0; b < v.n_bodies; b++) { if (members[b] != v.body_count[b]) { return false; } }
        return true;
    }

    // one array: read into dst, or skipped when dst is null
    bool readArray(std::FILE* f, void* dst, size_t bytes)
    {
        if (!bytes) { return true; }
        if (!dst) { return std::fseek(f, static_cast<long>(bytes), SEEK_CUR) == 0; }
        return std::fread(dst, 1, bytes, f) == bytes;
    }

    // one array: written from src, or as zeros when src is null
    bool writeArray(std::FILE* f, const void* src, size_t bytes)
    {
        if (!bytes) { return true; }
        if (src) { return std::fwrite(src, 1, bytes, f) == bytes; }
        std::vector<char> z(std::min<size_t>(bytes, size_t(1) << 20), 0);
        for (size_t done = 0; done < bytes;)
        {
            const size_t n = std::min(z.size(), bytes - done);
            if (std::fwrite(z.data(), 1, n, f) != n) { return false; }
            done += n;
        }
        return true;
    }
}

extern "C" int tsp_state_file_info(const char* path, uint32_t counts[4], uint64_t* step, double* time)
{
    if (!path || !counts) { return TSP_ERR_INVALID; }
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) { return TSP_ERR_INVALID; }
    Header h{};
    const int rc = readHeader(fh.f, h);
    if (rc != TSP_OK) { return rc; }
    counts[0] = static_cast<uint32_t>(h.nF); counts[1] = static_cast<uint32_t>(h.nB);
    counts[2] = static_cast<uint32_t>(h.nR); counts[3] = static_cast<uint32_t>(h.nRB);
    if (step) { *step = h.step; }
    if (time) { *time = h.time; }
    return TSP_OK;
}

extern "C" int tsp_state_file_read(const char* path, tsp_state_view* v)
{
    if (!path || !v) { return TSP_ERR_INVALID; }
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) { return TSP_ERR_INVALID; }
    Header h{};
    const int rc = readHeader(fh.f, h);
    if (rc != TSP_OK) { return rc; }
    if (v->n_fluid != h.nF || v->n_boundary != h.nB || v->n_rigid != h.nR || v->n_bodies != h.nRB) { return TSP_ERR_INVALID; }
    std::FILE* f = fh.f;
    bool ok = readArray(f, v->fluid_id, 8ull * h.nF);
    for (int q = 0; q < TSP_F_COUNT; q++) { ok = ok && readArray(f, v->fluid[q], 4ull * h.nF); }
    ok = ok && readArray(f, v->boundary_id, 8ull * h.nB);
    for (int q = 0; q < TSP_B_COUNT; q++) { ok = ok && readArray(f, v->boundary[q], 4ull * h.nB); }
    ok = ok && readArray(f, v->rigid_id, 8ull * h.nR);
    ok = ok && readArray(f, v->rigid_body, 4ull * h.nR);
    for (int q = 0; q < TSP_R_COUNT; q++) { ok = ok && readArray(f, v->rigid[q], 4ull * h.nR); }
    ok = ok && readArray(f, v->body_id, 4ull * h.nRB);
    ok = ok && readArray(f, v->body_count, 4ull * h.nRB);
    for (int q = 0; q < TSP_RB_COUNT; q++) { ok = ok && readArray(f, v->body[q], 4ull * h.nRB); }
    if (!ok || !validContents(*v)) { return TSP_ERR_INVALID; }
    v->step = h.step;
    v->time = h.time;
    return TSP_OK;
}

extern "C" int tsp_state_file_write(const char* path, const tsp_state_view* v)
{
    if (!path || !v) { return TSP_ERR_INVALID; }
    if (!std::isfinite(v->time) || v->time < 0.0 || v->step == std::numeric_limits<uint64_t>::max() || !validContents(*v))
    {
        return TSP_ERR_INVALID;
    }
    // written beside the destination and renamed over it, so a reader never sees a partial file
    // (SaveOptions::atomicWrite with tempSuffix ".tmp", serialization.h:33-34)
    const std::string tmp = std::string(path) + ".tmp";
    bool ok = false;
    {
        File fh;
        fh.f = std::fopen(tmp.c_str(), "wb");
        if (!fh.f) { return TSP_ERR_INVALID; }
        Header h{};
        std::memcpy(h.magic, kMagic, 8);
        h.endianTag = kEndian;
        h.headerBytes = sizeof(Header);
        h.step = v->step;
        h.time = v->time;
        h.nF = v->n_fluid; h.nB = v->n_boundary; h.nR = v->n_rigid; h.nRB = v->n_bodies;
        std::FILE* f = fh.f;
        ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
        ok = ok && writeArray(f, v->fluid_id, 8ull * h.nF);
        for (int q = 0; q < TSP_F_COUNT; q++) { ok = ok && writeArray(f, v->fluid[q], 4ull * h.nF); }
        ok = ok && writeArray(f, v->boundary_id, 8ull * h.nB);
        for (int q = 0; q < TSP_B_COUNT; q++) { ok = ok && writeArray(f, v->boundary[q], 4ull * h.nB); }
        ok = ok && writeArray(f, v->rigid_id, 8ull * h.nR);
        ok = ok && writeArray(f, v->rigid_body, 4ull * h.nR);
        for (int q = 0; q < TSP_R_COUNT; q++) { ok = ok && writeArray(f, v->rigid[q], 4ull * h.nR); }
        ok = ok && writeArray(f, v->body_id, 4ull * h.nRB);
        ok = ok && writeArray(f, v->body_count, 4ull * h.nRB);
        for (int q = 0; q < TSP_RB_COUNT; q++) { ok = ok && writeArray(f, v->body[q], 4ull * h.nRB); }
        ok = ok && std::fflush(f) == 0;
        ok = (std::fclose(fh.f) == 0) && ok;
        fh.f = nullptr;
    }
    if (ok && std::rename(tmp.c_str(), path) == 0) { return TSP_OK; }
    std::remove(tmp.c_str());
    return TSP_ERR_INVALID;
}
