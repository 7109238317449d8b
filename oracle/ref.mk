# Builds oracle/_ref/libdortref.so: dort's own pt-path sources, its PLY reader
# (with the rply C parser it links), image I/O (stb) and convergence test, compiled
# where they lie under /root/reference, linked with ref_driver.cpp and
# ref_io_driver.cpp (TEST INFRASTRUCTURE ONLY; the output stays out of git).
#
#   make -f oracle/ref.mk            (needs /root/reference; nothing is copied)
#
# dort's Tupfile.lua builds with clang++ and its "fast" flags (-O3 -DNDEBUG
# -DDORT_DISABLE_STAT); the compiler matters: clang evaluates call arguments left
# to right, which fixes the order of the two draws in Sampler::random_2d.  The
# sources predate C++17 and rely on <functional>, <string> and <stdexcept>
# arriving through other headers; -include supplies those standard headers.
REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
REF_CXX ?= $(firstword $(wildcard /opt/rocm/llvm/bin/clang++) clang++)
CXXFLAGS := -O3 -std=c++14 -fPIC -DDORT_DISABLE_STAT -DNDEBUG \
  -include functional -include string -include stdexcept \
  -I$(REF)/include -I$(REF)/extern -I$(REF)/extern/stb -I$(REF)/extern/rply -w
REF_CC ?= $(firstword $(wildcard /opt/rocm/llvm/bin/clang) clang)

SRCS := path_renderer renderer film bvh bvh_primitive mesh_bvh_primitive triangle \
  primitive scene bsdf lambert_material microfacet rough_dielectric_material \
  dielectric_material fresnel diffuse_light shape light environment_light \
  infinite_light piecewise_distrib_2d discrete_distrib_1d monte_carlo \
  pinhole_camera camera transform random_sampler sampler sampled_filter filter \
  stats thread_pool box geometry math image image_texture texture_maps \
  list_primitive mat triangle_shape triangle_shape_primitive mesh_triangle_primitive \
  ply_mesh convergence_test light_renderer
OBJS := $(addprefix $(OUT)/obj/,$(addsuffix .o,$(SRCS))) $(OUT)/obj/rply.o $(OUT)/obj/ref_io_driver.o

all: $(OUT)/libdortref.so

$(OUT)/obj/%.o: $(REF)/src/dort/%.cpp
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/ref_driver.o: $(HERE)ref_driver.cpp $(HERE)../include/dortgpu.h
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/ref_io_driver.o: $(HERE)ref_io_driver.cpp
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -c $< -o $@

# rply (extern/rply, the PLY parser dort links) is C
$(OUT)/obj/rply.o: $(REF)/extern/rply/rply.c
	@mkdir -p $(OUT)/obj
	$(REF_CC) -O2 -fPIC -w -I$(REF)/extern/rply -c $< -o $@

$(OUT)/libdortref.so: $(OBJS) $(OUT)/obj/ref_driver.o
	$(REF_CXX) -shared -o $@ $^ -lpthread

# dort with dort_amd::GpuPathRenderer (dort_amd/host) in its Renderer seam, over
# libdortgpu.so; built when that library exists (target `gpu`).
GPU_LIB := $(HERE)../dort_amd/libdortgpu.so
HOST := $(HERE)../dort_amd/host

$(OUT)/obj/gpu_path_renderer.o: $(HOST)/gpu_path_renderer.cpp $(HOST)/gpu_path_renderer.hpp $(HERE)../include/dortgpu.h
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -I$(HERE)../include -c $< -o $@

$(OUT)/obj/ref_gpu_glue.o: $(HERE)ref_gpu_glue.cpp $(HOST)/gpu_path_renderer.hpp $(HERE)../include/dortgpu.h
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -I$(HERE)../include -c $< -o $@

$(OUT)/libdortref_gpu.so: $(OBJS) $(OUT)/obj/ref_driver.o $(OUT)/obj/gpu_path_renderer.o $(OUT)/obj/ref_gpu_glue.o $(GPU_LIB)
	$(REF_CXX) -shared -o $@ $(filter %.o,$^) -L$(dir $(GPU_LIB)) -ldortgpu -Wl,-rpath,'$$ORIGIN/../../dort_amd' -lpthread

gpu: $(OUT)/libdortref_gpu.so

clean:
	rm -rf $(OUT)

.PHONY: all clean gpu
