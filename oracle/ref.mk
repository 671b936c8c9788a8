# Compile the reference's own single-rank FP64 sources that build without
# generated headers or MPI, straight from the reference tree, into
# oracle/_ref/libref_<nd0>x<nd1>x<nd2>x<nd3>.so (one library per lattice, as
# the reference fixes the lattice at compile time).  Test infrastructure only.
REF     ?= /root/reference/src
OUT     := _ref
CC      ?= gcc
LATTICES ?= 4x4x4x4 8x8x8x8 4x6x8x10
SRCS := OpenAcc/geometry.c OpenAcc/backfield.c OpenAcc/fermionic_utilities.c \
        RationalApprox/rationalapprox.c Mpi/multidev.c OpenAcc/stouting.c \
        OpenAcc/su3_utilities.c OpenAcc/plaquettes.c OpenAcc/su3_measurements.c \
        OpenAcc/rectangles.c OpenAcc/md_parameters.c Meas/baryon_number_utilities.c
TILES := -DDEODOEGANG3=8 -DDEODOETILE0=8 -DDEODOETILE1=8 -DDEODOETILE2=8 \
         -DSTAPGANG3=8 -DSTAPTILE0=8 -DSTAPTILE1=8 -DSTAPTILE2=8 \
         -DIMPSTAPGANG3=8 -DIMPSTAPTILE0=8 -DIMPSTAPTILE1=8 -DIMPSTAPTILE2=8 \
         -DSIGMAGANG3=8 -DSIGMATILE0=8 -DSIGMATILE1=8 -DSIGMATILE2=8
CFLAGS := -std=gnu99 -O2 -fPIC -fcommon -w -Wl,-Bsymbolic -DNRANKS_D3=1 -DNREPLICAS=1 -DACTION_TYPE=TLSM $(TILES) -I$(REF)

LIBS := $(foreach L,$(LATTICES),$(OUT)/libref_$(L).so)

all: $(LIBS)

dims = $(subst x, ,$(1))
$(OUT)/libref_%.so: ref_harness.c
	@mkdir -p $(OUT)
	$(CC) $(CFLAGS) -DLOC_N0=$(word 1,$(call dims,$*)) -DLOC_N1=$(word 2,$(call dims,$*)) \
	  -DLOC_N2=$(word 3,$(call dims,$*)) -DLOC_N3=$(word 4,$(call dims,$*)) -shared -o $@ \
	  ref_harness.c $(addprefix $(REF)/,$(SRCS)) -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
