// examples/PathTracer/main.cpp written against include/wayland_hip.hpp: the
// same sequence of ContextManager, build inputs, StaticAccelStructure,
// Module + IdentifyPrograms, Pipeline, SBT, Launcher, denoiser and
// FromFloatToChar, on a mesh dump instead of the assimp import.
//
//   example_pathtracer <mesh dump> <program source> <width> <height> <spp> <maxDepth> <out>
//
// Built twice: as an executable, and as a shared library exporting
// example_pathtracer_main so a process that already drives the GPU can run it
// in-process.
//
// Mesh dump (little endian): u32 M, then per mesh u32 nv, u32 nt, f32 kd[3],
// f32 ke[3], f32 vertices[3 nv], f32 normals[3 nv], u32 indices[3 nt].
// Output: f32 colour[4 W H], then u8 bytes[4 W H] from FromFloatToChar.
#include "wayland_hip.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>

using namespace Wayland;

namespace
{
struct MeshDump
{
    std::vector<float> vertices, normals;
    std::vector<unsigned int> indices;
    float kd[3], ke[3];
};

std::vector<MeshDump> read_meshes(const char *path)
{
    std::ifstream f(path, std::ios::binary);
    HostUtils::CheckError((bool)f, "cannot open the mesh dump");
    auto u32 = [&] {
        uint32_t v = 0;
        f.read(reinterpret_cast<char *>(&v), 4);
        return v;
    };
    std::vector<MeshDump> meshes(u32());
    for (auto &m : meshes)
    {
        const uint32_t nv = u32(), nt = u32();
        f.read(reinterpret_cast<char *>(m.kd), 12);
        f.read(reinterpret_cast<char *>(m.ke), 12);
        m.vertices.resize(3 * (size_t)nv);
        m.normals.resize(3 * (size_t)nv);
        m.indices.resize(3 * (size_t)nt);
        f.read(reinterpret_cast<char *>(m.vertices.data()), 12 * (std::streamsize)nv);
        f.read(reinterpret_cast<char *>(m.normals.data()), 12 * (std::streamsize)nv);
        f.read(reinterpret_cast<char *>(m.indices.data()), 12 * (std::streamsize)nt);
    }
    HostUtils::CheckError((bool)f, "truncated mesh dump");
    return meshes;
}

template<typename T> T *to_device(const std::vector<T> &v)
{
    T *d = nullptr;
    HostUtils::CheckHip(hipMalloc((void **)&d, v.size() * sizeof(T)), "hipMalloc");
    HostUtils::CheckHip(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice),
                        "hipMemcpy");
    return d;
}
} // namespace

extern "C" int example_pathtracer_main(int argc, char **argv)
{
    if (argc != 8)
    {
        std::fprintf(stderr, "usage: %s <mesh dump> <program source> <width> <height> <spp> "
                             "<maxDepth> <out>\n", argv[0]);
        return 2;
    }
    try
    {
        Optix::ContextManager manager;
        manager.SetCachePath(".");
        const auto meshes = read_meshes(argv[1]);
        const unsigned int width = (unsigned)std::atoi(argv[3]), height = (unsigned)std::atoi(argv[4]);
        const unsigned int spp = (unsigned)std::atoi(argv[5]), maxDepth = (unsigned)std::atoi(argv[6]);

        Optix::TriangleBuildInputArray arr{ meshes.size() };
        for (const auto &m : meshes)
            arr.AddBuildInput({ std::span<const float>{ m.vertices } },
                              std::span<const unsigned int>{ m.indices },
                              Optix::GeometryFlags::DiableAnyHit);
        Optix::StaticAccelStructure as{ arr };

        auto pipelineConfig = Optix::PipelineConfig::GetDefault().SetNumPayloadValues(5);
        Optix::Module myModule{ argv[2], Optix::ModuleConfig::GetDefaultRef(), pipelineConfig };
        Optix::ProgramGroupArray groups{ 3 };
        myModule.IdentifyPrograms({ argv[2] }, groups);
        Optix::Pipeline pipeline{ groups, maxDepth + 1, as.GetDepth(), 2, 2, pipelineConfig };

        // GenerateHitData (main.cpp:12-33): one record per build input
        std::vector<void *> device_buffers;
        std::vector<Optix::SBTData<wl_example_hit_data>> hits(meshes.size());
        std::vector<std::size_t> hitIdx(meshes.size(), groups.find("__closesthit__PT"));
        for (std::size_t i = 0; i < meshes.size(); ++i)
        {
            std::vector<int32_t> ind(meshes[i].indices.begin(), meshes[i].indices.end());
            int32_t *d_ind = to_device(ind);
            float *d_nrm = to_device(meshes[i].normals);
            device_buffers.push_back(d_ind);
            device_buffers.push_back(d_nrm);
            hits[i].data = wl_example_hit_data{ { meshes[i].kd[0], meshes[i].kd[1], meshes[i].kd[2] },
                                                { meshes[i].ke[0], meshes[i].ke[1], meshes[i].ke[2] },
                                                d_ind, d_nrm };
        }
        Optix::SBTData<void> raygen, miss;
        Optix::ShaderBindingTable sbt{ raygen, 0, miss, 1, hits, hitIdx.data(), groups };

        float *color = nullptr;
        HostUtils::CheckHip(hipMalloc((void **)&color, 16 * (size_t)width * height), "hipMalloc");
        // Example::Camera{ {-0.23, 2.58, 5}, up {0, 1, 0}, gaze {0, 0, -1} }.ToDeviceCamera(1):
        // the ratio is depth * tanf(fov / 2) with fov = 45 taken as radians
        const float vRatio = 1.0f * std::tan(45.0f / 2);
        wl_example_params param{};
        param.sampleNum = spp;
        param.maxDepth = maxDepth;
        param.camera = wl_example_camera{ 1.0f * vRatio, vRatio, { -0.23f, 2.58f, 5.0f },
                                          { 0.0f, 0.0f, -1.0f }, { 0.0f, 1.0f, 0.0f } };
        param.colorBuffer = color;
        param.traversable = as.GetHandle();
        Optix::Launcher launcher{ param };
        launcher.Launch(pipeline, nullptr, sbt, width, height);
        HostUtils::CheckHip(hipDeviceSynchronize(), "hipDeviceSynchronize");

        try
        {
            Optix::LDRDenoiser denoiser{};
            denoiser.Denoise();
        }
        catch (const HostUtils::Error &e)
        {
            std::printf("denoiser: %s\n", e.what());
        }

        auto byteResult = FromFloatToChar(color, 4 * (std::size_t)width * height);
        std::vector<float> host(4 * (std::size_t)width * height);
        HostUtils::CheckHip(hipMemcpy(host.data(), color, host.size() * 4, hipMemcpyDeviceToHost),
                            "hipMemcpy");
        std::ofstream out(argv[7], std::ios::binary);
        out.write(reinterpret_cast<const char *>(host.data()), (std::streamsize)host.size() * 4);
        out.write(reinterpret_cast<const char *>(byteResult.data()),
                  (std::streamsize)byteResult.size());
        HostUtils::CheckError((bool)out, "cannot write the output");
        (void)hipFree(color);
        for (void *p : device_buffers)
            (void)hipFree(p);
        std::printf("rendered %ux%u, %u spp\n", width, height, spp);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

#ifdef WL_EXAMPLE_EXECUTABLE
int main(int argc, char **argv) { return example_pathtracer_main(argc, argv); }
#endif
