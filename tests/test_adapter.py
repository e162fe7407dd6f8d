"""The drop-in for TracerCUDA (meturay_amd/adapter/TracerMRH.cpp) without a
GPU: libTracerMRH.so exports the factory pair of TracerCUDA/EntryPoint.h:12-16
and implements MRay's own TracerSystemI / GPUTracerI / GPUSceneI, compiled
from the reference's headers.  tests/cpp/mray_tracer_thread drives it
through those headers as MRay's TracerThread does and checks the factory,
Initialize's error without a device, the tracers' option sets and errors
(PathTracer.cu:120-142, WFPGTracer.cu:1355-1435), the filter options
(TracerSystemCUDA.cpp:109-141) and work without a loaded scene."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADAPTER = os.path.join(ROOT, "meturay_amd", "adapter", "_build", "libTracerMRH.so")
HOST = os.path.join(ROOT, "tests", "cpp", "_build", "mray_tracer_thread")
SCENE = os.path.join(ROOT, "tests", "golden", "scenes", "room.json")

built = pytest.mark.skipif(not (os.path.exists(ADAPTER) and os.path.exists(HOST)),
                           reason="the adapter is built only where the reference's headers are "
                                  "(make -C meturay_amd/adapter all host)")


@built
def test_adapter_exports_the_reference_factory():
    out = subprocess.run(["nm", "-D", "--defined-only", ADAPTER], capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert {"GenerateTracerSystem", "DeleteTracerSystem"} <= syms
    # the vtables implement the reference's own interface classes
    assert "_ZTI13TracerSystemI" in syms
    lib = os.path.join(ROOT, "meturay_amd", "libmrh.so")
    core = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    assert "GenerateTracerSystem" not in core      # one factory, in the drop-in only


@built
def test_tracer_thread_protocol_without_gpu():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")   # no device, even where one exists
    r = subprocess.run([HOST, "cpu", ADAPTER, SCENE], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    assert sum(line.startswith("ok ") for line in r.stdout.splitlines()) >= 25
