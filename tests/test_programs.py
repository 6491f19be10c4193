"""The OptiX program interface: program discovery by name (CPU), and launches
of the two built-in programs through pipeline / SBT / LaunchParams with the
reference's record layouts (GPU), checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
from wayland_amd import scenes

PT_SOURCE = r'''
extern "C" __constant__ Programs::PathTracing::LaunchParams param;
/* PG id - 0 */
extern "C" __global__ void __raygen__RenderFrame()
{ }
/* PG id - 1 */
extern "C" __global__ void __miss__radiance() {}
/* PG id - 2 */
extern "C" __global__ void __miss__shadow() {}
/* PG id - 3 */
extern "C" __global__ void __closesthit__radiance() {}
/* PG id - 4 */
extern "C" __global__ void __anyhit__shadow() {}
'''

EXAMPLE_SOURCE = r'''
extern "C" __constant__ LaunchParam param;
extern "C" __global__ void __raygen__RenderFrame()
{ }
extern "C" __global__ void __miss__Empty() { }
extern "C" __global__ void __closesthit__PT() { }
'''


def test_identify_pathtracing_program_ids():
    import wayland_amd as wl
    g, _, _ = wl.identify_programs(PT_SOURCE)
    assert [k for k, _ in g] == [0, 1, 1, 2, 2]
    assert g[0][1][0] == "__raygen__RenderFrame"
    assert g[1][1][0] == "__miss__radiance" and g[2][1][0] == "__miss__shadow"
    assert g[3][1] == ("", "", "__closesthit__radiance")
    assert g[4][1] == ("", "__anyhit__shadow", "")


def test_identify_example_program_ids():
    import wayland_amd as wl
    g, _, _ = wl.identify_programs(EXAMPLE_SOURCE)
    assert [(k, e[0] or e[2]) for k, e in g] == [
        (0, "__raygen__RenderFrame"), (1, "__miss__Empty"), (2, "__closesthit__PT")]


def test_identify_rules():
    import wayland_amd as wl
    src = '''
    extern "C" __global__ void __closesthit__radiance() {}
    extern  "C"   __global__
       void __miss__a () {}
    extern "C" __global__ void __miss__spaced( ) {}
    extern "C" __global__ void __raygen__x() {}
    extern "C" __global__ void __anyhit__radiance() {}
    extern "C" __global__ void __intersection__sphere() {}
    extern "C" __global__ void __direct_callable__f() {}
    extern "C" __global__ void __exception__e() {}
    extern "C" __global__ void __raygen__withargs(int) {}
    extern "C" __global__ void notaprogram() {}
    extern "C" __device__ void __raygen__device() {}
    '''
    g, _, _ = wl.identify_programs(src)
    # raygen/miss in file order, then hit groups in order of first appearance;
    # same-suffix closest-hit and any-hit share a group; "( )" is not "()"
    assert g[0] == (1, ("__miss__a", "", ""))
    assert g[1] == (0, ("__raygen__x", "", ""))
    assert g[2] == (2, ("", "__anyhit__radiance", "__closesthit__radiance"))
    assert g[3] == (2, ("__intersection__sphere", "", ""))
    assert len(g) == 4


@pytest.mark.gpu
def test_unknown_program_set_rejected(gpu_ctx):
    import wayland_amd as wl
    with pytest.raises(wl.WlError) as e:
        wl.Pipeline(gpu_ctx, 'extern "C" __global__ void __raygen__Other() {}')
    assert e.value.status == 7


def _tan22_5():
    return float(np.tan(np.float32(22.5)))


@pytest.mark.gpu
def test_example_pathtracer_launch_matches_oracle(gpu_ctx):
    """examples/PathTracer: cornell.obj (the reference's data file), Lambertian
    recursive PT through pipeline + SBT + LaunchParam, vs the oracle's
    restatement of PathTracer.cu."""
    import os
    import torch
    import wayland_amd as wl
    from wayland_amd import _lib, obj
    print("case: examples/PathTracer cornell.obj 48x48, 8 spp, maxDepth 9", flush=True)
    path = os.path.join(os.path.dirname(__file__), "golden", "cornell.obj")
    meshes, mats = obj.load_obj(path)
    assert sum(m.indices.shape[0] for m in meshes) == 32
    sc = scenes.Scene([scenes.Mesh(m.vertices, m.indices, m.normals) for m in meshes])
    accel = wl.Accel(gpu_ctx, sc)
    pipe = wl.Pipeline(gpu_ctx, EXAMPLE_SOURCE, max_trace_depth=10)
    assert pipe.builtin == 2
    keep = []
    hit = []
    for m in meshes:
        ind = torch.from_numpy(m.indices.astype(np.int32)).cuda()
        nrm = torch.from_numpy(m.normals).cuda()
        keep += [ind, nrm]
        d = _lib.ExampleHitData()
        d.color[:] = [float(x) for x in mats[m.material].kd]
        d.emission[:] = [float(x) for x in mats[m.material].ke]
        d.indices = ind.data_ptr()
        d.normals = nrm.data_ptr()
        hit.append(pipe.record(pipe.group_index("__closesthit__PT"), d))
    sbt = wl.SBT(pipe, pipe.record(0), [pipe.record(1)], hit)
    W = H = 48
    spp, depth = 8, 9
    buf = torch.zeros((W * H, 4), dtype=torch.float32, device="cuda")
    p = _lib.ExampleParams()
    p.sampleNum, p.maxDepth = spp, depth
    v = _tan22_5()
    p.camera.horizontalRatio, p.camera.verticalRatio = v, v
    p.camera.position[:] = [-0.23, 2.58, 5.0]
    p.camera.gaze[:] = [0.0, 0.0, -1.0]
    p.camera.up[:] = [0.0, 1.0, 0.0]
    p.colorBuffer = buf.data_ptr()
    p.traversable = wl.traversable(accel)
    wl.launch(pipe, p, sbt, W, H)
    torch.cuda.synchronize()
    g = buf.cpu().numpy()
    opt = ob.OraclePT(sc)
    ec = ob.ExampleConfig(W, H, spp, depth)
    ec.cam_pos[:] = list(p.camera.position)
    ec.cam_gaze[:] = list(p.camera.gaze)
    ec.cam_up[:] = list(p.camera.up)
    ec.horizontal_ratio, ec.vertical_ratio = v, v
    mc = np.stack([mats[m.material].kd for m in meshes])
    me = np.stack([mats[m.material].ke for m in meshes])
    o = opt.example_render(mc, me, ec, 0, W)
    err = np.abs(g - o) / np.maximum(np.abs(o), 1e-4)
    frac = float((err[:, :3].max(1) <= 1e-5).mean())
    exact = float((g.view(np.uint32) == o.view(np.uint32)).all(-1).mean())
    print(f"pixels within 1e-5: {frac:.5f}, bit-exact {exact:.5f}; "
          f"mean {g[:, :3].mean():.6f} vs {o[:, :3].mean():.6f}")
    # a pixel averages 8 paths of up to 10 bounces; one path moved by an ulp
    # of the device sinf/cosf onto another triangle changes the pixel
    # (an emitter of radiance 60 makes such a pixel differ by O(1), so image
    # means are not a sharp check here; the per-pixel fractions are)
    assert frac >= 0.99 and exact >= 0.97
    assert (g[:, 3] == 1).all()
    # FromFloatToChar
    u8 = wl.float_to_u8(buf.reshape(-1)).cpu().numpy()
    ref = (np.minimum(g.reshape(-1), 1.0) * 255).astype(np.uint8)
    assert np.array_equal(u8, ref)


@pytest.mark.gpu
def test_pathtracing_program_launch_matches_oracle(gpu_ctx):
    """src/Frontend/Programs/PathTracing through pipeline + SBT (HitData with
    device vertex/index/normal pointers, two ray types) + LaunchParams with a
    device DeviceAreaLight table: frames 0..2 vs the oracle's frames and
    progressive film."""
    import torch
    import wayland_amd as wl
    from wayland_amd import _lib
    print("case: PathTracing program via wl_launch, cornell 32x32, 3 frames", flush=True)
    W = H = 32
    sc = scenes.cornell_box(W, H)
    accel = wl.Accel(gpu_ctx, sc)
    pipe = wl.Pipeline(gpu_ctx, PT_SOURCE)
    assert pipe.builtin == 1
    keep, hit, lights = [], [], []
    for i, m in enumerate(sc.meshes):
        ind = torch.from_numpy(m.indices.astype(np.int32)).cuda()
        ver = torch.from_numpy(m.vertices).cuda()
        nrm = torch.from_numpy(m.normals).cuda()
        keep += [ind, ver, nrm]
        d = _lib.PathTracingHitData()
        d.hasNormal = 1
        d.meshID = i
        d.areaLightID = len(lights) if m.light_L is not None else 0xFFFFFFFF
        d.disneyMat[:] = [float(x) for x in m.material[:15]]
        d.indices, d.vertices, d.normals = ind.data_ptr(), ver.data_ptr(), nrm.data_ptr()
        hit.append(pipe.record(3, d))
        hit.append(pipe.record(4, _lib.PathTracingHitData()))
        if m.light_L is not None:
            L = _lib.DeviceAreaLight()
            L.twoSided, L.hasNormal, L.triangleNum = 0, 1, m.indices.shape[0]
            L.L[:] = list(m.light_L)
            L.vertices, L.normals, L.indices = ver.data_ptr(), nrm.data_ptr(), ind.data_ptr()
            lights.append(L)
    larr = (_lib.DeviceAreaLight * len(lights))(*lights)
    dl = torch.from_numpy(np.frombuffer(bytes(larr), np.uint8).copy()).cuda()
    miss = C.c_float * 3
    sbt = wl.SBT(pipe, pipe.record(0), [pipe.record(1, miss(0, 0, 0)),
                                        pipe.record(2, miss(0, 0, 0))], hit)
    rad = torch.zeros((W * H, 4), dtype=torch.float32, device="cuda")
    col = torch.zeros((W * H, 4), dtype=torch.uint8, device="cuda")
    p = _lib.PathTracingParams()
    p.fbSize[:] = [W, H]
    p.radianceBuffer, p.colorBuffer = rad.data_ptr(), col.data_ptr()
    cam = sc.camera
    p.camera.pos[:], p.camera.lookat[:] = list(cam.pos), list(cam.lookat)
    p.camera.up[:], p.camera.right[:] = list(cam.up), list(cam.right)
    p.areaLightCount, p.areaLights = len(lights), dl.data_ptr()
    p.traversable = wl.traversable(accel)
    opt = ob.OraclePT(sc)
    oacc = np.zeros((W * H, 4), np.float32)
    ocol = np.zeros((W * H, 4), np.uint8)
    for f in range(3):
        p.frameID = f
        wl.launch(pipe, p, sbt, W, H)
        o, _ = opt.render_frame(opt.config(W, H), f)
        ob.film_update(oacc, ocol, o, f)
    torch.cuda.synchronize()
    a = rad.cpu().numpy()
    close = np.abs(a - oacc) <= 1e-5 * np.maximum(np.abs(oacc), 1e-3)
    assert close.all(-1).mean() >= 0.999
    c = col.cpu().numpy()
    assert (np.abs(c.astype(int) - ocol.astype(int)) <= 1).all(-1).mean() >= 0.999


@pytest.mark.gpu
def test_pathtracing_program_with_texture_objects(gpu_ctx):
    """Hit records with hasTexture = 1, a wl_texture_create handle and device
    texcoords (PathTracingProgramManager.cpp:97-119) render the oracle's
    textured frames."""
    import torch
    import wayland_amd as wl
    from wayland_amd import _lib
    print("case: PathTracing program with texture objects, cornell 32x32, 2 frames", flush=True)
    W = H = 32
    sc = scenes.textured_cornell_box(W, H)
    accel = wl.Accel(gpu_ctx, sc)
    pipe = wl.Pipeline(gpu_ctx, PT_SOURCE)
    texs = [wl.Texture(gpu_ctx, img) for img in sc.textures]
    keep, hit, lights = [], [], []
    for i, m in enumerate(sc.meshes):
        ind = torch.from_numpy(m.indices.astype(np.int32)).cuda()
        ver = torch.from_numpy(m.vertices).cuda()
        nrm = torch.from_numpy(m.normals).cuda()
        keep += [ind, ver, nrm]
        d = _lib.PathTracingHitData()
        d.hasNormal = 1
        d.meshID = i
        d.areaLightID = len(lights) if m.light_L is not None else 0xFFFFFFFF
        d.disneyMat[:] = [float(x) for x in m.material[:15]]
        d.indices, d.vertices, d.normals = ind.data_ptr(), ver.data_ptr(), nrm.data_ptr()
        if m.texture is not None:
            d.hasTexture = 1
            d.texture = texs[m.texture].handle
            if m.uv is not None:
                uv = torch.from_numpy(np.ascontiguousarray(m.uv, np.float32)).cuda()
                keep.append(uv)
                d.texcoords = uv.data_ptr()
        hit.append(pipe.record(3, d))
        hit.append(pipe.record(4, _lib.PathTracingHitData()))
        if m.light_L is not None:
            L = _lib.DeviceAreaLight()
            L.twoSided, L.hasNormal, L.triangleNum = 0, 1, m.indices.shape[0]
            L.L[:] = list(m.light_L)
            L.vertices, L.normals, L.indices = ver.data_ptr(), nrm.data_ptr(), ind.data_ptr()
            lights.append(L)
    larr = (_lib.DeviceAreaLight * len(lights))(*lights)
    dl = torch.from_numpy(np.frombuffer(bytes(larr), np.uint8).copy()).cuda()
    miss = C.c_float * 3
    sbt = wl.SBT(pipe, pipe.record(0), [pipe.record(1, miss(0, 0, 0)),
                                        pipe.record(2, miss(0, 0, 0))], hit)
    rad = torch.zeros((W * H, 4), dtype=torch.float32, device="cuda")
    col = torch.zeros((W * H, 4), dtype=torch.uint8, device="cuda")
    p = _lib.PathTracingParams()
    p.fbSize[:] = [W, H]
    p.radianceBuffer, p.colorBuffer = rad.data_ptr(), col.data_ptr()
    cam = sc.camera
    p.camera.pos[:], p.camera.lookat[:] = list(cam.pos), list(cam.lookat)
    p.camera.up[:], p.camera.right[:] = list(cam.up), list(cam.right)
    p.areaLightCount, p.areaLights = len(lights), dl.data_ptr()
    p.traversable = wl.traversable(accel)
    opt = ob.OraclePT(sc)
    oacc = np.zeros((W * H, 4), np.float32)
    ocol = np.zeros((W * H, 4), np.uint8)
    for f in range(2):
        p.frameID = f
        wl.launch(pipe, p, sbt, W, H)
        o, _ = opt.render_frame(opt.config(W, H), f)
        ob.film_update(oacc, ocol, o, f)
    torch.cuda.synchronize()
    for t in texs:
        t.close()
    a = rad.cpu().numpy()
    close = np.abs(a - oacc) <= 1e-5 * np.maximum(np.abs(oacc), 1e-3)
    assert close.all(-1).mean() >= 0.999
