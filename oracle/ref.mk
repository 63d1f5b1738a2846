# Builds the reference's own Host library (sjseo00/dsDecompX, unmodified, the
# nine sources its CMakeLists.txt:86-95 lists, HAS_CUDA=0 as its CMake sets for
# a build without CUDA) straight from /root/reference with the system g++ and
# MPICH from /opt/conda (its headers, and libmpi.so named by path so that the
# system libstdc++ is linked; run the program with
# LD_LIBRARY_PATH=/usr/lib/x86_64-linux-gnu:/opt/conda/lib), plus this
# project's fixture writer
# (tests/golden/ref_grid_fixture.cpp) linked against it.  TEST
# INFRASTRUCTURE: outputs go to oracle/_ref/ only; nothing of it is part of
# the product, and the GPU runs never need it (the fixtures it writes are
# committed under tests/golden/).  Built only where /root/reference exists.
#
#   make -f oracle/ref.mk            (from the repository root)
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
REF ?= /root/reference
CXX := g++
MPI ?= /opt/conda
OUT := $(HERE)_ref
# the reference Host build: g++ -O3, x86-64 without FMA (its CMake adds no -march)
FLAGS := -O3 -std=c++17 -DHAS_CUDA=0 -DHAS_NCCL=0 -DHAS_NVSHMEM=0 -DENABLE_PERF_COUNTERS=0 \
         -I$(REF)/include -I$(REF)/src -isystem $(MPI)/include
SRCS := backend_selection.cpp decomp.cpp runtime.cpp topology.cpp grid.cpp halo.cpp migration.cpp \
        buffer.cpp backends/mpi_backend.cpp
OBJS := $(addprefix $(OUT)/,$(subst /,_,$(SRCS:.cpp=.o)))

all: $(OUT)/ref_grid_fixture

$(OUT):
	mkdir -p $(OUT)

define compile_rule
$(OUT)/$(subst /,_,$(1:.cpp=.o)): $(REF)/src/$(1) | $(OUT)
	$(CXX) $(FLAGS) -c $$< -o $$@
endef
$(foreach s,$(SRCS),$(eval $(call compile_rule,$(s))))

$(OUT)/libdsdecompx_ref.a: $(OBJS)
	ar rcs $@ $^

$(OUT)/ref_grid_fixture: $(HERE)../tests/golden/ref_grid_fixture.cpp $(OUT)/libdsdecompx_ref.a
	$(CXX) $(FLAGS) $< $(OUT)/libdsdecompx_ref.a $(MPI)/lib/libmpi.so -o $@

.PHONY: all
