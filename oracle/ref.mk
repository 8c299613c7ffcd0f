# Builds the reference's own randombenchmark (the CPU baseline beside the GPU
# numbers) from the unmodified sources of the CAPS tree, out of tree: the
# objects and the binary go to oracle/_ref/ only, nothing is copied.
#   make -f oracle/ref.mk REF=/root/reference
# MPI: the image's MPICH wrapper with the system g++; BLAS dgemm_: MKL's
# single dynamic library with GNU OpenMP threading.  NUM_THREADS is the
# compile-time OpenMP thread count of the reference (multiply.cpp:11,
# make.incs/make.inc.linux:10).
REF ?= /root/reference
OUT ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
THREADS ?= 16
CXX := MPICH_CXX=g++ /opt/conda/bin/mpicxx
# the reference defines an unguarded two-argument max macro (library.h:1); the
# C++ headers it reaches later through math.h are included first instead
CXXFLAGS := -O3 -fopenmp -DNUM_THREADS=$(THREADS) -w -include cmath -include limits \
            -include algorithm
LIBS := -L/opt/conda/lib -Wl,-rpath,/opt/conda/lib -lmkl_rt
SRCS := randombenchmark.cpp communication.cpp matrix.cpp multiply.cpp memory.cpp \
        dgemm-blas.cpp counters.cpp library.cpp summa1d.cpp

$(OUT)/randombenchmark: $(addprefix $(REF)/,$(SRCS))
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF) $^ -o $@ $(LIBS)

.PHONY: clean
clean:
	rm -f $(OUT)/randombenchmark
