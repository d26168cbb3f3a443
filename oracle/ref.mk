# ref.mk — builds the REFERENCE solver core (JaiVerma20/ps26119-solver, read-only under
# /root/reference) straight from its own C++ sources with g++, into oracle/_ref/.
#
# This is test infrastructure only: the result is a CPU checker for the oracle restatement in
# oracle/ and an optional CPU baseline ("reference" kind). Nothing in the product links it.
# No reference build system is run and no reference file is copied: the sources are compiled
# where they lie. oracle/ref_driver.cpp (ours) is linked in to expose a small extern "C"
# surface (full solves with presolve on/off, scaling factors, ‖Ã‖₂, KKT, CPU backend ops,
# certified bound, solution check) for ctypes.
#
#   make -f oracle/ref.mk            (from the repo root; needs /root/reference)
REF      ?= /root/reference
OUT      := oracle/_ref
CXX      ?= g++
# Same semantics as the reference's Release build on x86-64 (no FMA in the baseline ISA);
# contraction is switched off explicitly so the bits do not depend on -march.
CXXFLAGS := -std=c++20 -O2 -fPIC -ffp-contract=off -DNDEBUG -DPS26119_HAVE_MPS_READER=1 \
            -I$(REF)/include -I$(REF)/src -w
SRCS := src/core/model.cpp src/core/status.cpp src/core/solve.cpp src/core/c_api.cpp \
        src/core/safe_bound.cpp src/core/implied_bounds.cpp src/core/presolve.cpp \
        src/io/lpm_reader.cpp src/io/solution_writer.cpp src/io/solution_reader.cpp \
        src/io/mps_reader.cpp src/oracle/dense_simplex.cpp src/oracle/dense_tableau.cpp \
        src/core/solution_checker.cpp src/core/certificates.cpp src/core/ranging.cpp \
        src/la/csr.cpp src/la/csc.cpp src/la/sparse_lu.cpp src/simplex/simplex_scaling.cpp \
        src/simplex/primal_simplex.cpp src/pdhg/scaling.cpp src/pdhg/cpu_backend.cpp \
        src/pdhg/engine.cpp src/pdhg/pdlp.cpp src/pdhg/r2hpdhg.cpp src/pdhg/batch.cpp \
        src/mip/branch_and_bound.cpp
OBJS := $(patsubst src/%.cpp,$(OUT)/obj/%.o,$(SRCS))

# The same core with the reference's GPU hook switched on (only src/pdhg/cpu_backend.cpp reads
# PS26119_HAVE_CUDA) and answered by the MI355X backend adapter: the reference's drivers over
# the psh_* kernels, for the GPU tests. Needs libps26119_hip.so (make -C ps26119-solver_amd/csrc).
PKG      := ps26119-solver_amd/csrc
HIP_LIB  := $(PKG)/hip/libps26119_hip.so
HIP_OBJS := $(filter-out $(OUT)/obj/pdhg/cpu_backend.o,$(OBJS)) $(OUT)/obj_hip/cpu_backend.o \
            $(OUT)/obj_hip/hip_backend.o $(OUT)/obj_hip/ref_hip_hook.o $(OUT)/obj/ref_driver.o

all: $(OUT)/libps26119_ref.so $(OUT)/libps26119_ref_hip.so

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/ref_driver.o: oracle/ref_driver.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libps26119_ref.so: $(OBJS) $(OUT)/obj/ref_driver.o
	$(CXX) -shared -o $@ $^ -lpthread

$(OUT)/obj_hip/cpu_backend.o: $(REF)/src/pdhg/cpu_backend.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -DPS26119_HAVE_CUDA -c $< -o $@

$(OUT)/obj_hip/hip_backend.o: $(PKG)/adapter/hip_backend.cpp include/ps26119_hip.h
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -DPS26119_HAVE_CUDA -Iinclude -Wall -Wextra -c $< -o $@

$(OUT)/obj_hip/ref_hip_hook.o: oracle/ref_hip_hook.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -DPS26119_HAVE_CUDA -c $< -o $@

$(OUT)/libps26119_ref_hip.so: $(HIP_OBJS) $(HIP_LIB)
	$(CXX) -shared -o $@ $(HIP_OBJS) -L$(PKG)/hip -lps26119_hip \
	    -Wl,-rpath,'$$ORIGIN/../../$(PKG)/hip' -lpthread

clean:
	rm -rf $(OUT)

.PHONY: all clean
