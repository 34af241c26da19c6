# Compiles the reference's own pure-Fortran element layer, from its sources
# where they lie under /root/reference, into oracle/_ref/libhr_ref.so.
# Only ROCm flang is needed; nothing is copied out of /root/reference and all
# outputs (objects, .mod files, the library) go to oracle/_ref/.
# The PETSc-coupled layers (assembly, RK4 drivers) need PETSc/MPI/HDF5, which
# this image lacks, and are not built.
FLANG   ?= /opt/rocm/lib/llvm/bin/flang
REFDIR  ?= /root/reference/step_4_3D_Forward_and_Inverse
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
FFLAGS  ?= -O0 -fPIC
SRCS    := Parameters Shape_Functions Source1 PML Element_Matrices_3D
OBJS    := $(addprefix $(OUT)/,$(addsuffix .o,$(SRCS)))

all: $(OUT)/libhr_ref.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/Parameters.o: $(REFDIR)/Parameters.F90 | $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o $@

$(OUT)/%.o: $(REFDIR)/%.F90 $(OUT)/Parameters.o | $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o $@

$(OUT)/ref_shim.o: $(HERE)ref_shim.f90 $(OUT)/Parameters.o | $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o $@

$(OUT)/libhr_ref.so: $(OBJS) $(OUT)/ref_shim.o
	$(FLANG) -shared -o $@ $^

clean:
	rm -rf $(OUT)

.PHONY: all clean
