"""The reference-side binding that INTEGRATION.md shows a maintainer adding
(the C++ block under "Reference-side binding") compiles against
include/heipa.h: the block is extracted, its reference includes are replaced
by stand-ins with the shapes the binding uses (Kokkos views, the Graph and
Partition members it reads), and g++ type-checks explicit instantiations of
every replaced function.  The stand-ins only give the binding's own calls into
the C-ABI something to bind to; they are not the reference's types."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = r'''
#include <cstdint>
#include <cstdlib>
#include <utility>
#include "heipa.h"
using u32 = uint32_t;
using weight_t = int32_t;
using partition_t = uint32_t;
using vertex_t = uint32_t;
namespace Kokkos {
template<typename T, typename S = void> struct View;
template<typename T, typename S> struct View<T *, S> {
    T *p = nullptr;
    size_t n = 0;
    View() = default;
    View(const char *, size_t k) : n(k) {}
    T *data() const { return p; }
    size_t extent(int) const { return n; }
    T &operator()(size_t i) const { return p[i]; }
};
template<typename E> struct RangePolicy { RangePolicy(E &, size_t, size_t) {} };
template<typename P, typename F> void parallel_for(const char *, const P &, const F &) {}
}
#define KOKKOS_LAMBDA [=]
struct DeviceMemorySpace {};
struct DeviceExecutionSpace { void fence() {} };
struct KokkosMemoryStack {};
using UnmanagedDeviceWeight = Kokkos::View<weight_t *>;
namespace GPU_HeiPa {
struct Graph {
    vertex_t n, m;
    int64_t g_weight;
    bool uniform_vertex_weights, uniform_edge_weights;
    Kokkos::View<uint32_t *> neighborhood, edges_v;
    Kokkos::View<int32_t *> edges_w, weights;
};
struct Partition {
    Kokkos::View<partition_t *> map;
    UnmanagedDeviceWeight bweights;
};
}
'''

INSTANTIATE = r'''
struct Oracle { Kokkos::View<int32_t *> w_mtx; };
template weight_t GPU_HeiPa::edge_cut<false>(const GPU_HeiPa::Graph &, const GPU_HeiPa::Partition &, DeviceExecutionSpace &);
template std::pair<weight_t, weight_t> GPU_HeiPa::jet_refine<false, false>(GPU_HeiPa::Graph &, GPU_HeiPa::Partition &,
    partition_t, weight_t, bool, u32, weight_t, weight_t, KokkosMemoryStack &, DeviceExecutionSpace &);
template std::pair<weight_t, weight_t> GPU_HeiPa::ProMap_jet_refine<false, false, Oracle>(GPU_HeiPa::Graph &,
    GPU_HeiPa::Partition &, Oracle &, partition_t, weight_t, u32, weight_t, weight_t, KokkosMemoryStack &,
    DeviceExecutionSpace &);
'''


def _binding_block():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    part = text.split("## Reference-side binding", 1)[1]
    m = re.search(r"```cpp\n(.*?)```", part, re.S)
    assert m, "no C++ block under the binding heading"
    return "\n".join(l for l in m.group(1).splitlines() if not l.startswith("#include"))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_binding_compiles_against_the_header(tmp_path):
    src = tmp_path / "binding.cpp"
    src.write_text(PRELUDE + _binding_block() + INSTANTIATE)
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
