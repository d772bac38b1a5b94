# Test infrastructure only: builds the reference CPU code (1-rank oracle) from
# its own sources where they lie, plus the dump hook in ref_hook.c.
#
#   make -f oracle/ref.mk            -> oracle/_ref/mumatscale_ref
#
# Compile definitions follow the reference's CMakeLists.txt:64-69 (INDEX_SEP,
# NUC_SEP, EBUG=0; no GPU offload) at its default Release -O3, with gcc at the
# default x86-64 -march (no FMA).  MPI (MPICH) and HDF5/zlib are the copies
# installed in the image under /opt/conda; nothing is stubbed.  Build outputs
# go to oracle/_ref/ only.  ref_hook.c wraps doiteration (state dumps) and
# writeMainCheckpoint (skipped on request: it crashes before any checkpoint
# file is complete).

REF_SRC ?= /root/reference/src
CONDA   ?= /opt/conda
HERE    := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)/_ref
OBJ     := $(OUT)/obj

SRCS := packing calculate ca_main checkpoint curvature distribute face_util \
        file_io globals growth functions grain io_vtk io_xdmf ll loop profiler \
        stopwatch read_ctrl setup temperature vtkCompress

CC      := gcc
CFLAGS  := -std=gnu99 -O3 -DNDEBUG -fopenmp -DINDEX_SEP -DNUC_SEP -DEBUG=0 \
           -I$(REF_SRC) -I$(CONDA)/include -w
LDFLAGS := -L$(CONDA)/lib -Wl,-rpath,$(CONDA)/lib -Wl,--wrap=doiteration -Wl,--wrap=writeMainCheckpoint
LIBS    := -lhdf5 -lmpi -lz -lm -fopenmp

OBJS := $(patsubst %,$(OBJ)/%.o,$(SRCS)) $(OBJ)/ref_hook.o

all: $(OUT)/mumatscale_ref

$(OBJ):
	mkdir -p $(OBJ)

$(OBJ)/%.o: $(REF_SRC)/%.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@

$(OBJ)/ref_hook.o: $(HERE)/ref_hook.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/mumatscale_ref: $(OBJS)
	$(CC) $(OBJS) -o $@ $(LDFLAGS) $(LIBS)

clean:
	rm -rf $(OUT)

.PHONY: all clean
