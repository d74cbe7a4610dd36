# Builds the reference e8qlib (E8Q + ME8Q CPU library) from its own sources
# where they lie, into oracle/_ref/ only.  Used as the checker that pins the
# oracle restatement and as the "reference" CPU baseline.  Nothing under
# oracle/_ref/ is committed.
#
# Flags follow the reference CMakeLists.txt:9-38,144-149 (C++17 ISO mode,
# -O3, -fno-math-errno, OpenMP, E8Q_ISA=avx2 = its default backend), with
# -march=native replaced by -march=x86-64-v3 (AVX2+FMA) so the same library
# runs on this host and on the GPU host.  `-include cstddef`: the reference
# rotator.hpp uses size_t without including <cstddef>, which GCC 11's
# <cstdint> does not pull in (a standard header, forced in, no stand-in).
REF      ?= /root/reference
OUT      ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX      ?= g++
CXXFLAGS := -include cstddef -std=c++17 -O3 -march=x86-64-v3 -fno-math-errno -fopenmp -fPIC \
            -DE8Q_HAVE_AVX2 -I$(REF)/include -I$(REF)/src
AVX2FLAGS := -mavx2 -mfma -mpopcnt

GENERIC := utils/cpu_features utils/rotator quantization/e8q_codec \
           quantization/me8q_plan quantization/me8q_codec index/ivf_route \
           index/e8q_scan index/e8q_build index/e8q_io index/e8q_search \
           index/e8q_oracle index/me8q_common index/me8q_io index/me8q_measure \
           index/me8q_build index/me8q_search index/me8q_fast simd/dispatch \
           simd/kernels_generic simd/kernels_dispatch simd/scan_kernels_generic \
           simd/scan_kernels_dispatch simd/fastscan_generic simd/fastscan_dispatch \
           simd/me8q_score_generic simd/me8q_score_dispatch simd/e8q_search_generic \
           simd/e8q_search_dispatch simd/me8q_prune_generic simd/me8q_prune_dispatch
AVX2    := simd/kernels_avx2 simd/scan_kernels_avx2 simd/fastscan_avx2 \
           simd/me8q_score_avx2 simd/e8q_search_avx2 simd/me8q_prune_avx2

GEN_OBJ  := $(addprefix $(OUT)/obj/,$(addsuffix .o,$(GENERIC)))
AVX2_OBJ := $(addprefix $(OUT)/obj/,$(addsuffix .o,$(AVX2)))

all: $(OUT)/libe8qref.so

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(AVX2_OBJ): CXXFLAGS += $(AVX2FLAGS)

$(OUT)/libe8qref.so: $(GEN_OBJ) $(AVX2_OBJ)
	$(CXX) -shared -fopenmp -o $@ $^

.PHONY: all

# Golden-vector writer (tests/golden/gen_golden.cpp) linked against the
# reference library; output stays in oracle/_ref/.
$(OUT)/gen_golden: $(dir $(lastword $(MAKEFILE_LIST)))../tests/golden/gen_golden.cpp $(OUT)/libe8qref.so
	$(CXX) -include cstddef -std=c++17 -O2 -I$(REF)/include $< -o $@ -L$(OUT) -le8qref -Wl,-rpath,'$$ORIGIN'

all: $(OUT)/gen_golden

# CPU baseline timer over the reference library (oracle/ref_bench.cpp).
$(OUT)/ref_bench: $(dir $(lastword $(MAKEFILE_LIST)))ref_bench.cpp $(OUT)/libe8qref.so
	$(CXX) -include cstddef -std=c++17 -O2 -fopenmp -I$(REF)/include $< -o $@ -L$(OUT) -le8qref -Wl,-rpath,'$$ORIGIN'

all: $(OUT)/ref_bench
