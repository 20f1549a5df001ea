"""The reference-side binding (include/swarm_crowdsim_adapter.hpp) compiles against the reference's own headers.

The adapter keeps class CrowdSim's surface (include/swarm/crowd_sim.hpp:53-194) over the C ABI and converts the
reference's GpuBoid / GpuGridInfo / GpuCellData records (include/swarm/gpu_types.hpp:17-36). Here a translation unit
that uses every method is compiled with the reference's include/ and its vendored glm, then linked against
libswarm_hip.so (nothing runs: no GPU here). Skipped where the reference's tree is absent.
"""
import os
import subprocess

import pytest

import swarm_hip as sh

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

USE = r"""
#include "swarm_crowdsim_adapter.hpp"
int main() {
    swarm_scene_desc d{};
    swarm_config cfg{ 0.5f, 20.0f, 32, 1 };
    swarm::HipCrowdSimAdapter sim(d, 256, cfg, 0.25f / 31.25f, 2.0f * 17.490671f / 31.25f);
    std::vector<swarm::GpuBoid> b(256);
    std::vector<float> r(256, 0.0064f), v0(256, 0.75f);
    sim.upload(b, r, v0);
    sim.setProjectionIterations(32);
    sim.step();
    sim.stepExitChoice();
    const std::vector<swarm::GpuBoid> s = sim.currentState(), p = sim.previousState();
    const std::vector<swarm::GpuGridInfo> g = sim.grid();
    const std::vector<swarm::GpuCellData> c = sim.cells();
    const std::vector<uint32_t> a = sim.assignedExits();
    return int(s.size() + p.size() + g.size() + c.size() + a.size()) == 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include", "swarm")), reason="the reference's tree is absent")
def test_adapter_compiles_and_links_against_the_reference_headers(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(USE)
    exe = tmp_path / "use"
    lib_dir = os.path.dirname(sh.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O0", "-I", os.path.join(REF, "include"), "-I", os.path.join(REF, "external", "glm"),
           "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe), "-L", lib_dir, "-lswarm_hip",
           "-Wl,--allow-shlib-undefined", "-Wl,-rpath," + lib_dir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
