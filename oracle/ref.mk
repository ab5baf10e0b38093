# Builds oracle/_ref/libfsibe_ref.so from the reference's own Fortran sources,
# compiled where they lie (REF, read-only) with amdflang, plus the driver
# oracle/ref_harness.f90.  Only routines that compile unmodified from their own
# files are included; the FFT backends (FFTW/MKL absent, cof_temp.f90 needs
# share_vars symbols that no longer exist), create_mask.f90 (Intel-only
# logical ==) and Interpolation.f90 (needs nrutil, which flang rejects) are
# unbuildable here and are not stubbed.
#
# oracle/_ref/save_gif_ref: the reference's SaveGIF (FieldExport.f90) and
# GIF writer (gif_util.f90), compiled where they lie, driven by
# oracle/ref_gif_driver.f90.  FieldExport.f90 also holds SaveField, which
# calls BeamIO's WriteHeaderToFile; BeamIO needs the MKL solid solver and is
# not built, and the driver never calls SaveField, so the link leaves that
# one symbol unresolved instead of supplying it.
#
# Usage: make -f oracle/ref.mk   (from the repository root)
REF ?= /root/reference
FLANG ?= /opt/rocm/bin/amdflang
FFLAGS ?= -O2 -fPIC -fdefault-real-8 -fdefault-double-8
OUT := oracle/_ref
SRCS := share_vars cofdx cofdy cofdxdx cofdydy cofdxdy poisson cal_vis \
        dealiase_mask mean_velocity mouvement integrate_position
OBJS := $(addprefix $(OUT)/,$(addsuffix .o,$(SRCS)))

all: $(OUT)/libfsibe_ref.so $(OUT)/save_gif_ref

$(OUT)/%.o: $(REF)/%.f90 | $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o $(notdir $@)

$(OUT)/cofdx.o $(OUT)/cofdy.o $(OUT)/cofdxdx.o $(OUT)/cofdydy.o $(OUT)/cofdxdy.o \
$(OUT)/poisson.o $(OUT)/cal_vis.o $(OUT)/dealiase_mask.o $(OUT)/mean_velocity.o \
$(OUT)/mouvement.o: $(OUT)/share_vars.o
$(OUT)/integrate_position.o: $(OUT)/mouvement.o

$(OUT)/ref_harness.o: oracle/ref_harness.f90 $(OUT)/mouvement.o | $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $(abspath oracle/ref_harness.f90) -o ref_harness.o

$(OUT)/libfsibe_ref.so: $(OBJS) $(OUT)/ref_harness.o
	$(FLANG) -shared -o $@ $^

$(OUT)/gif_util.o: $(REF)/gif_util.f90 | $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o gif_util.o
$(OUT)/FieldExport.o: $(REF)/FieldExport.f90 $(OUT)/share_vars.o $(OUT)/gif_util.o | $(OUT)
	cd $(OUT) && $(FLANG) $(FFLAGS) -c $< -o FieldExport.o

$(OUT)/save_gif_ref: oracle/ref_gif_driver.f90 $(OUT)/FieldExport.o $(OUT)/gif_util.o $(OUT)/share_vars.o
	cd $(OUT) && $(FLANG) $(FFLAGS) -o save_gif_ref $(abspath oracle/ref_gif_driver.f90) FieldExport.o \
	    gif_util.o share_vars.o -Wl,--unresolved-symbols=ignore-in-object-files

$(OUT):
	mkdir -p $(OUT)

clean:
	rm -rf $(OUT)

.PHONY: all clean
