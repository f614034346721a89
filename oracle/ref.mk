# Builds the reference's own CPU programs from their Fortran sources where they lie under
# $(REF), into oracle/_ref/ only (test infrastructure: the CPU baseline, and a checker of the
# oracle on decks no shipped curve covers).  Nothing from the reference is copied into the
# repository, and none of its shipped objects, modules or executables is used: every object
# is compiled here from the .f90 sources, in a build directory of our own.
#
#   make -f oracle/ref.mk            (needs the reference sources, amdflang and MKL)
#
# The reference Makefiles (sim-trhepd-rheed/src/Makefile:7-14, str-lapack/src/Makefile:7-13)
# ask for gfortran and static MKL; this image has neither, so the same sources are compiled
# with ROCm's Flang (amdflang, OpenMP on) and linked against the image's MKL shared libraries
# (LP64 interface, GNU threading layer).  Programs:
#   surf.exe        sim-trhepd-rheed surface program (int_method 1-8 from stdin: "dz method")
#   bulk.exe        sim-trhepd-rheed bulk program (writes bulkP.b)
#   surf-lapack.exe str-lapack transfer-matrix surface program (stdin: "dz")
#   libh211b.so     the step-size controller h211b_control.f90 alone (functions h211b_limit_,
#                   h211b_filter_; no program calls it), for checking its restatements
REF ?= /root/reference
FLANG ?= /opt/rocm/bin/amdflang
MKLDIR ?= /opt/conda/lib
FFLAGS ?= -O3 -fopenmp
HERE := $(abspath $(dir $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)/_ref
MKL := -L$(MKLDIR) -Wl,-rpath,$(MKLDIR) -lmkl_intel_lp64 -lmkl_gnu_thread -lmkl_core

SIM := $(REF)/sim-trhepd-rheed/src
LAP := $(REF)/str-lapack/src
# module producers before their users (surf_prkn / surf_magnus / surf_rk need matcomp,
# h211b_control and surfio; surflf needs all of them)
SIM_SURF := h211b_control matcomp surfio matgs surf_prkn surf_magnus surf_rk surflf surfsub domainsum surfm
SIM_BULK := bulkio bulk bulksub bulkm
COMMON := asf strfac scpot trmatg gcmi
LAP_SURF := surfio surf surfsub domainsum surfm

all: $(OUT)/surf.exe $(OUT)/bulk.exe $(OUT)/surf-lapack.exe $(OUT)/libh211b.so

# one build directory per program family, so that each program's modules are its own
define compile
	@mkdir -p $(1)
	cd $(1) && for f in $(3); do $(FLANG) $(FFLAGS) -c $(2)/$$f.f90 -o $$f.o || exit 1; done
endef

$(OUT)/surf.exe: $(foreach f,$(SIM_SURF) $(COMMON),$(SIM)/$(f).f90)
	$(call compile,$(OUT)/sim-surf,$(SIM),$(SIM_SURF) $(COMMON))
	cd $(OUT)/sim-surf && $(FLANG) $(FFLAGS) -o $@ $(addsuffix .o,$(SIM_SURF) $(COMMON)) $(MKL)

$(OUT)/bulk.exe: $(foreach f,$(SIM_BULK) $(COMMON),$(SIM)/$(f).f90)
	$(call compile,$(OUT)/sim-bulk,$(SIM),$(SIM_BULK) $(COMMON))
	cd $(OUT)/sim-bulk && $(FLANG) $(FFLAGS) -o $@ $(addsuffix .o,$(SIM_BULK) $(COMMON)) $(MKL)

$(OUT)/surf-lapack.exe: $(foreach f,$(LAP_SURF) $(COMMON),$(LAP)/$(f).f90)
	$(call compile,$(OUT)/lapack-surf,$(LAP),$(LAP_SURF) $(COMMON))
	cd $(OUT)/lapack-surf && $(FLANG) $(FFLAGS) -o $@ $(addsuffix .o,$(LAP_SURF) $(COMMON)) $(MKL)

$(OUT)/libh211b.so: $(SIM)/h211b_control.f90
	@mkdir -p $(OUT)/h211b
	cd $(OUT)/h211b && $(FLANG) -O3 -fPIC -shared $(SIM)/h211b_control.f90 -o $@

clean:
	rm -rf $(OUT)
.PHONY: all clean
