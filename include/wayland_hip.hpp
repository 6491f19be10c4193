/*
 * wayland_hip.hpp — C++ face of the C ABI in wayland_hip.h, with the class
 * names and call shapes of the reference's OptiX wrapper
 * (src/Backend/Optix/Core), so host code written against it, such as
 * examples/PathTracer/main.cpp, keeps its structure:
 *
 *   Wayland::Optix::ContextManager      ContextManager.h:43-57
 *   TriangleBuildInputArray             BuildInput.h:154-198
 *   StaticAccelStructure                AccelStructure.h:80-86, Traversable.h:59-78
 *   Module, ModuleConfig, PipelineConfig Module.h:46-140, ProgramConfig.h:16-101
 *   ProgramGroupArray                   ProgramGroup.h:13-39
 *   Pipeline                            Pipeline.h:15-60
 *   SBTData<T>                          SBTData.h:8-24
 *   ShaderBindingTable                  ShaderBindingTable.h:13-80
 *   Launcher                            Launcher.h:12-60
 *   LDRDenoiser                         Denoiser.h (not available: Denoise() throws)
 *   FromFloatToChar                     examples/PathTracer/Conversion.h:5
 *
 * Differences, each forced by the backend:
 *  - Modules are not compiled: kernels are built ahead of time for gfx950.
 *    A Module keeps its source text, and IdentifyPrograms reads the program
 *    names from it (wl_identify_programs).  SetOptixSDKPath and AddArgs
 *    are accepted and do nothing.
 *  - One device per ContextManager (the first visible one, or the first
 *    id given).  Objects created while it lives use it, as the reference's
 *    current-context setter does (ContextManager.h:84-123).
 *  - SBT record headers are packed for the Pipeline built from the same
 *    ProgramGroupArray, so the Pipeline must exist before the SBT (it does in
 *    every reference program).
 *  - Failures throw Wayland::HostUtils::Error, carrying the wl_status and
 *    the library's message.  The reference throws only in its debug builds
 *    (ErrorCheck.cpp:14-22).
 */
#ifndef WAYLAND_HIP_HPP
#define WAYLAND_HIP_HPP

#include "wayland_hip.h"

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <ranges>
#include <span>
#include <sstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

namespace Wayland
{
namespace HostUtils
{
/// A failed library call: status plus the library's message.
class Error : public std::runtime_error
{
public:
    Error(wl_status status, const std::string &what)
        : std::runtime_error(what), status_(status)
    {
    }
    wl_status status() const noexcept { return status_; }

private:
    wl_status status_;
};

inline void Check(wl_status s, const char *call)
{
    if (s != WL_SUCCESS)
    {
        const char *msg = wl_last_error_message();
        throw Error(s, std::string(call) + ": " + wl_status_string(s) + (msg ? ": " : "") +
                           (msg ? msg : ""));
    }
}

/// ErrorCheck.h CheckError(bool, msg).
inline void CheckError(bool ok, const char *msg)
{
    if (!ok)
        throw Error(WL_ERROR_INVALID_VALUE, msg);
}

inline void CheckHip(hipError_t e, const char *call)
{
    if (e != hipSuccess)
        throw Error(WL_ERROR_DEVICE, std::string(call) + ": " + hipGetErrorString(e));
}
} // namespace HostUtils

namespace Optix
{
namespace detail
{
inline wl_context &current_context()
{
    static thread_local wl_context ctx = nullptr;
    return ctx;
}
inline wl_context require_context()
{
    wl_context c = current_context();
    if (!c)
        throw HostUtils::Error(WL_ERROR_INVALID_VALUE, "no live ContextManager on this thread");
    return c;
}
inline std::string read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw HostUtils::Error(WL_ERROR_INVALID_VALUE, "cannot read " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}
} // namespace detail

/// One device and its stream; current for this thread while it lives.
class ContextManager
{
public:
    ContextManager() : ContextManager(std::span<const int>{}) {}
    explicit ContextManager(std::span<const int> visibleDevices)
    {
        const int dev = visibleDevices.empty() ? 0 : visibleDevices[0];
        HostUtils::Check(wl_context_create(dev, &ctx_), "wl_context_create");
        prev_ = detail::current_context();
        detail::current_context() = ctx_;
    }
    ContextManager(const ContextManager &) = delete;
    ContextManager &operator=(const ContextManager &) = delete;
    ContextManager(ContextManager &&o) noexcept
        : ctx_(std::exchange(o.ctx_, nullptr)), prev_(o.prev_)
    {
        if (detail::current_context() == ctx_)
            detail::current_context() = ctx_;
    }
    ~ContextManager()
    {
        if (!ctx_)
            return;
        if (detail::current_context() == ctx_)
            detail::current_context() = prev_;
        wl_context_destroy(ctx_);
    }
    void SetCachePath(std::string_view path)
    {
        HostUtils::Check(wl_set_cache_path(ctx_, std::string(path).c_str()), "wl_set_cache_path");
    }
    wl_context Get() const noexcept { return ctx_; }
    void *GetStream() const
    {
        void *s = nullptr;
        HostUtils::Check(wl_context_get_stream(ctx_, &s), "wl_context_get_stream");
        return s;
    }

private:
    wl_context ctx_ = nullptr;
    wl_context prev_ = nullptr;
};

/// OPTIX_GEOMETRY_FLAG_* values (BuildInput.h:50-56).
enum class GeometryFlags : unsigned int
{
    None = 0,
    DiableAnyHit = 1,
    SingleAnyhit = 2,
    FaceCulling = 4
};

/// OPTIX_BUILD_FLAG_* values (AccelStructure.h:12-21); the BVH8 build
/// ignores them.
enum class BuildFlags : unsigned int
{
    None = 0,
    Update = 1,
    Compact = 2,
    FastTrace = 4,
    FastBuild = 8,
    VertexRandomAccess = 16
};

/// Triangle build inputs; host data is copied, as TriangleDataBuffer does.
class TriangleBuildInputArray
{
public:
    TriangleBuildInputArray() = default;
    explicit TriangleBuildInputArray(std::size_t expectBuildInputNum)
    {
        vertices_.reserve(expectBuildInputNum);
        triangles_.reserve(expectBuildInputNum);
        flags_.reserve(expectBuildInputNum);
    }

    /// vertices: one span per motion key (only one key here).
    void AddBuildInput(const std::vector<std::span<const float>> &vertices,
                       std::span<const unsigned int> triangles,
                       GeometryFlags flag = GeometryFlags::None)
    {
        if (vertices.size() != 1)
            throw HostUtils::Error(WL_ERROR_NOT_SUPPORTED, "motion keys are not supported");
        if (vertices[0].size() % 3 != 0 || triangles.size() % 3 != 0)
            throw HostUtils::Error(WL_ERROR_INVALID_VALUE, "vertices / triangles not in triples");
        vertices_.emplace_back(vertices[0].begin(), vertices[0].end());
        triangles_.emplace_back(triangles.begin(), triangles.end());
        flags_.push_back(static_cast<uint32_t>(flag));
    }
    void RemoveBuildInput(std::size_t idx) noexcept
    {
        if (idx >= vertices_.size())
            return;
        vertices_.erase(vertices_.begin() + (std::ptrdiff_t)idx);
        triangles_.erase(triangles_.begin() + (std::ptrdiff_t)idx);
        flags_.erase(flags_.begin() + (std::ptrdiff_t)idx);
    }
    unsigned int GetDepth() const noexcept { return 1; }
    std::size_t size() const noexcept { return vertices_.size(); }

    std::vector<wl_tri_input> inputs() const
    {
        std::vector<wl_tri_input> in(vertices_.size());
        for (std::size_t i = 0; i < in.size(); ++i)
            in[i] = wl_tri_input{ vertices_[i].data(), (uint32_t)(vertices_[i].size() / 3),
                                  triangles_[i].data(), (uint32_t)(triangles_[i].size() / 3),
                                  flags_[i], 1u };
        return in;
    }

private:
    std::vector<std::vector<float>> vertices_;
    std::vector<std::vector<unsigned int>> triangles_;
    std::vector<uint32_t> flags_;
};

/// The acceleration structure over a build input array.
class StaticAccelStructure
{
public:
    explicit StaticAccelStructure(const TriangleBuildInputArray &arr,
                                  BuildFlags flag = BuildFlags::None)
    {
        auto in = arr.inputs();
        HostUtils::Check(wl_accel_build_triangles(detail::require_context(), in.data(),
                                                  (uint32_t)in.size(),
                                                  static_cast<uint32_t>(flag), &a_),
                         "wl_accel_build_triangles");
    }
    StaticAccelStructure(const StaticAccelStructure &) = delete;
    StaticAccelStructure &operator=(const StaticAccelStructure &) = delete;
    StaticAccelStructure(StaticAccelStructure &&o) noexcept : a_(std::exchange(o.a_, nullptr)) {}
    ~StaticAccelStructure()
    {
        if (a_)
            wl_accel_destroy(a_);
    }
    /// The value a LaunchParams `traversable` field holds.
    uint64_t GetHandle() const
    {
        uint64_t h = 0;
        HostUtils::Check(wl_accel_traversable(a_, &h), "wl_accel_traversable");
        return h;
    }
    unsigned int GetDepth() const noexcept { return 1; }
    wl_accel Get() const noexcept { return a_; }

private:
    wl_accel a_ = nullptr;
};

/// Compile options of the reference; kept for call compatibility.
class ModuleConfig
{
public:
    static const ModuleConfig &GetDefaultRef() noexcept
    {
        static const ModuleConfig c;
        return c;
    }
    static ModuleConfig GetDefault() noexcept { return GetDefaultRef(); }
};

class PipelineConfig
{
public:
    static const PipelineConfig &GetDefaultRef() noexcept
    {
        static const PipelineConfig c;
        return c;
    }
    static PipelineConfig GetDefault() noexcept { return GetDefaultRef(); }
    PipelineConfig &SetNumPayloadValues(int n) noexcept
    {
        payload_values_ = n;
        return *this;
    }
    PipelineConfig &SetNumAttributeValues(int n) noexcept
    {
        attribute_values_ = n;
        return *this;
    }

private:
    int payload_values_ = 0, attribute_values_ = 2;
};

class Pipeline;

/// Program groups in IdentifyPrograms order (raygen, miss, hit groups).
class ProgramGroupArray
{
public:
    ProgramGroupArray() = default;
    explicit ProgramGroupArray(std::size_t expectedNum) { groups_.reserve(expectedNum); }
    ProgramGroupArray(const ProgramGroupArray &) = delete;
    ProgramGroupArray &operator=(const ProgramGroupArray &) = delete;
    ProgramGroupArray(ProgramGroupArray &&) noexcept = default;
    ProgramGroupArray &operator=(ProgramGroupArray &&) noexcept = default;

    std::size_t size() const noexcept { return groups_.size(); }
    const wl_program_group *data() const noexcept { return groups_.data(); }
    /// Index of the group holding entry point `name`, or size().
    std::size_t find(std::string_view name) const noexcept
    {
        for (std::size_t i = 0; i < groups_.size(); ++i)
            for (int k = 0; k < 3; ++k)
                if (name == groups_[i].entry[k])
                    return i;
        return groups_.size();
    }

private:
    friend class Module;
    friend class Pipeline;
    friend class ShaderBindingTable;
    std::vector<wl_program_group> groups_;
    mutable wl_pipeline pipeline_ = nullptr; // the pipeline built from these groups
};

/// A device-code module: its source text, from which program names are read.
class Module
{
public:
    explicit Module(std::string_view fileName,
                    const ModuleConfig & = ModuleConfig::GetDefaultRef(),
                    const PipelineConfig & = PipelineConfig::GetDefaultRef(),
                    const std::vector<std::string> & = {})
        : files_{ std::string(fileName) }
    {
    }
    explicit Module(const std::vector<std::string> &fileNames,
                    const ModuleConfig & = ModuleConfig::GetDefaultRef(),
                    const PipelineConfig & = PipelineConfig::GetDefaultRef(),
                    const std::vector<std::string> & = {})
        : files_(fileNames)
    {
    }
    template<typename T> static void SetOptixSDKPath(T &&) {}
    template<typename T> static void AddArgs(T &&) {}

    /// Module::IdentifyPrograms (Module-AutoConfig.cpp:76-178).
    void IdentifyPrograms(const std::vector<std::string> &identifyFiles, ProgramGroupArray &arr)
    {
        std::string src;
        for (const auto &f : identifyFiles)
            src += detail::read_file(f) + "\n";
        std::vector<wl_program_group> g(64);
        uint32_t n = 0;
        HostUtils::Check(wl_identify_programs(src.data(), src.size(), g.data(),
                                              (uint32_t)g.size(), &n),
                         "wl_identify_programs");
        g.resize(n);
        arr.groups_ = std::move(g);
    }

private:
    std::vector<std::string> files_;
};

/// The pipeline of a program set; recognises the set by its program names.
class Pipeline
{
public:
    Pipeline(const ProgramGroupArray &arr, unsigned int maxTraceDepth,
             unsigned int /*maxTraversableDepth*/, unsigned int /*maxCCDepth*/ = 0,
             unsigned int /*maxDCDepth*/ = 0,
             const PipelineConfig & = PipelineConfig::GetDefaultRef())
    {
        HostUtils::Check(wl_pipeline_create(detail::require_context(), arr.groups_.data(),
                                            (uint32_t)arr.groups_.size(), maxTraceDepth, &p_,
                                            &builtin_),
                         "wl_pipeline_create");
        arr.pipeline_ = p_;
    }
    Pipeline(const Pipeline &) = delete;
    Pipeline &operator=(const Pipeline &) = delete;
    Pipeline(Pipeline &&o) noexcept : p_(std::exchange(o.p_, nullptr)), builtin_(o.builtin_) {}
    ~Pipeline()
    {
        if (p_)
            wl_pipeline_destroy(p_);
    }
    wl_pipeline Get() const noexcept { return p_; }
    uint32_t builtin() const noexcept { return builtin_; }

private:
    wl_pipeline p_ = nullptr;
    uint32_t builtin_ = 0;
};

template<typename T>
    requires std::is_trivially_copyable_v<T> || std::is_same_v<T, void>
struct alignas(WL_SBT_RECORD_ALIGNMENT) SBTData
{
    std::byte header[WL_SBT_RECORD_HEADER_SIZE];
    T data;
    using value_type = T;
};

template<>
struct alignas(WL_SBT_RECORD_ALIGNMENT) SBTData<void>
{
    std::byte header[WL_SBT_RECORD_HEADER_SIZE];
    using value_type = void;
};

/// The SBT: packs each record's header for its program group, then copies
/// the records (ShaderBindingTable.h:13-80).
class ShaderBindingTable
{
public:
    template<typename RaygenData, typename MissRange, typename HitRange>
    ShaderBindingTable(SBTData<RaygenData> &raygen, std::size_t raygenIdx, MissRange &miss,
                       const std::size_t *missIdx, HitRange &hit, const std::size_t *hitIdx,
                       const ProgramGroupArray &arr)
    {
        auto *m = std::ranges::data(miss);
        auto *h = std::ranges::data(hit);
        const std::size_t nm = std::ranges::size(miss), nh = std::ranges::size(hit);
        pack(arr, &raygen, raygenIdx);
        for (std::size_t i = 0; i < nm; ++i)
            pack(arr, m + i, missIdx[i]);
        for (std::size_t i = 0; i < nh; ++i)
            pack(arr, h + i, hitIdx[i]);
        create(arr, &raygen, sizeof raygen, m, (uint32_t)nm, sizeof(*m), h, (uint32_t)nh,
               sizeof(*h));
    }
    template<typename RaygenData, typename MissData, typename HitRange>
    ShaderBindingTable(SBTData<RaygenData> &raygen, std::size_t raygenIdx,
                       SBTData<MissData> &miss, std::size_t missIdx, HitRange &hit,
                       const std::size_t *hitIdx, const ProgramGroupArray &arr)
    {
        auto *h = std::ranges::data(hit);
        const std::size_t nh = std::ranges::size(hit);
        pack(arr, &raygen, raygenIdx);
        pack(arr, &miss, missIdx);
        for (std::size_t i = 0; i < nh; ++i)
            pack(arr, h + i, hitIdx[i]);
        create(arr, &raygen, sizeof raygen, &miss, 1u, sizeof miss, h, (uint32_t)nh, sizeof(*h));
    }
    ShaderBindingTable(const ShaderBindingTable &) = delete;
    ShaderBindingTable &operator=(const ShaderBindingTable &) = delete;
    ShaderBindingTable(ShaderBindingTable &&o) noexcept : s_(std::exchange(o.s_, nullptr)) {}
    ~ShaderBindingTable()
    {
        if (s_)
            wl_sbt_destroy(s_);
    }
    wl_sbt Get() const noexcept { return s_; }

private:
    static wl_pipeline pipeline_of(const ProgramGroupArray &arr)
    {
        if (!arr.pipeline_)
            throw HostUtils::Error(WL_ERROR_INVALID_VALUE,
                                   "build the Pipeline of these program groups first");
        return arr.pipeline_;
    }
    static void pack(const ProgramGroupArray &arr, void *record, std::size_t idx)
    {
        HostUtils::Check(wl_sbt_pack_header(pipeline_of(arr), (uint32_t)idx, record),
                         "wl_sbt_pack_header");
    }
    void create(const ProgramGroupArray &arr, const void *rg, std::size_t rg_size,
                const void *miss, uint32_t nm, std::size_t ms, const void *hit, uint32_t nh,
                std::size_t hs)
    {
        HostUtils::Check(wl_sbt_create(pipeline_of(arr), rg, rg_size, miss, nm, ms, hit, nh, hs,
                                       &s_),
                         "wl_sbt_create");
    }
    wl_sbt s_ = nullptr;
};

/// Holds a copy of the launch parameters (Launcher.h:12-60).
class Launcher
{
public:
    template<typename T>
        requires std::is_trivially_copyable_v<std::remove_reference_t<T>>
    explicit Launcher(T &&data)
    {
        SetData(std::forward<T>(data));
    }
    Launcher(void *ptr, std::size_t size)
        : data_(static_cast<std::byte *>(ptr), static_cast<std::byte *>(ptr) + size)
    {
    }
    template<typename T>
        requires std::is_trivially_copyable_v<std::remove_reference_t<T>>
    void SetData(T &&data)
    {
        auto *p = reinterpret_cast<const std::byte *>(std::addressof(data));
        data_.assign(p, p + sizeof(std::remove_reference_t<T>));
    }
    void Launch(const Pipeline &pipeline, void *stream, const ShaderBindingTable &sbt,
                unsigned int width, unsigned int height, unsigned int depth = 1) const
    {
        HostUtils::Check(wl_launch(pipeline.Get(), stream, data_.data(), data_.size(), sbt.Get(),
                                   width, height, depth),
                         "wl_launch");
    }

private:
    std::vector<std::byte> data_;
};

/// The OptiX denoiser is a closed NVIDIA network with no counterpart here.
class LDRDenoiser
{
public:
    template<typename... A>
    explicit LDRDenoiser(A &&...)
    {
    }
    void Denoise()
    {
        throw HostUtils::Error(WL_ERROR_NOT_SUPPORTED,
                               "LDRDenoiser: no denoiser on this backend (the OptiX one is a "
                               "closed NVIDIA network)");
    }
};
} // namespace Optix
} // namespace Wayland

/// examples/PathTracer/Conversion.h: a device float buffer of `size` values
/// as host bytes, v > 1 -> 255, else (unsigned char)(v * 255).
inline std::vector<unsigned char> FromFloatToChar(float *d_src, std::size_t size)
{
    using Wayland::HostUtils::Check;
    using Wayland::HostUtils::CheckHip;
    std::vector<unsigned char> out(size);
    if (size == 0)
        return out;
    uint8_t *d_dst = nullptr;
    CheckHip(hipMalloc((void **)&d_dst, size), "hipMalloc");
    wl_status s = wl_float_to_u8(d_src, size, d_dst, nullptr);
    hipError_t e = s == WL_SUCCESS ? hipMemcpy(out.data(), d_dst, size, hipMemcpyDeviceToHost)
                                   : hipSuccess;
    (void)hipFree(d_dst);
    Check(s, "wl_float_to_u8");
    CheckHip(e, "hipMemcpy");
    return out;
}

#endif
