# Builds the REFERENCE implementation (daschaich/susy 4d_Q16, U(2) as its
# header fixes NCOL) from its sources where they lie, for use as a checker.
# Outputs go only to oracle/_ref/ (git-ignored):
#   oracle/_ref/libsusyref_U2.so  reference objects + oracle/ref_harness.c
#   oracle/_ref/susy_hmc_U2       the reference's own susy_hmc program
# LAPACK (zgetrf/zgetri/zheev, used by the reference's determinant.c and
# unit.c) comes from the image's MKL runtime library in /opt/conda/lib.
#
#   make -f oracle/ref.mk REF=/root/reference/4d_Q16 -j8
REF    ?= /root/reference/4d_Q16
OUT    := oracle/_ref
OBJ    := $(OUT)/obj
MKLDIR ?= /opt/conda/lib
CC     := gcc
CFLAGS := -O3 -fPIC -fno-strict-aliasing -w -DSINGLE -DPRECISION=2 \
          -D_FILE_OFFSET_BITS=64 -D_LARGEFILE_SOURCE64 \
          -DPHI_ALGORITHM -DHMC_ALGORITHM -I$(REF)/susy
LIBS   := -L$(MKLDIR) -lmkl_rt -Wl,-rpath,$(MKLDIR) -lm

SUSY_SRC := setup setup_lambda setup_offset action plaq ploop unit scalar_eig \
            link_trace widths determinant grsource utilities library_util \
            gauge_info congrad_multi update_o update_h setup_rhmc det_force
GEN_SRC  := io_helpers io_lat io_lat_utils make_lattice nersc_cksum ranstuff \
            gaugefix remap_stdio_from_args layout_hyper_prime io_ansi com_vanilla
LIB_SRC  := cmplx cexp clog csqrt ce_itheta dcmplx dcexp dclog dcsqrt \
            dce_itheta cmp_ahmat dump_ahmat make_ahmat uncmp_ahmat rand_ahmat \
            m_su2_mat_vec_n m_su2_mat_vec_a gaussrand z2rand byterevn m_mat_an \
            m_mat_na m_mat_nn realtr s_m_a_mat s_m_a_amat s_m_s_mat s_m_s_amat \
            adjoint mat_copy trace complextr dumpmat dumpsu2 clear_mat addmat \
            addamat submat subamat clearvec r_su2_hit_a l_su2_hit_n s_m_mat \
            s_m_amat s_a_d_mat cs_a_d_mat s_m_mat_nn s_m_mat_na s_m_mat_an \
            cs_m_mat cs_m_a_mat cs_m_a_amat cs_m_a_mata cs_m_s_mat

OBJS := $(SUSY_SRC:%=$(OBJ)/susy_%.o) $(GEN_SRC:%=$(OBJ)/gen_%.o) \
        $(LIB_SRC:%=$(OBJ)/lib_%.o)

all: $(OUT)/libsusyref_U2.so $(OUT)/susy_hmc_U2 $(OUT)/susy_refbench_U4 \
     $(OUT)/susy_hmc_U2_hip

$(OBJ):
	mkdir -p $(OBJ)

$(OBJ)/susy_%.o: $(REF)/susy/%.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@
$(OBJ)/gen_%.o: $(REF)/generic/%.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@
$(OBJ)/lib_%.o: $(REF)/libraries/%.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@
$(OBJ)/control.o: $(REF)/susy/control.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@
$(OBJ)/ref_harness.o: oracle/ref_harness.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/libsusyref_U2.so: $(OBJS) $(OBJ)/ref_harness.o
	$(CC) -shared -o $@ $^ $(LIBS)

$(OUT)/susy_hmc_U2: $(OBJS) $(OBJ)/control.o
	$(CC) -o $@ $^ $(LIBS)

# ---------------------------------------------------------------------------
# The drop-in test: the reference's susy_hmc with its own definitions of the
# five replaced symbols localised (objcopy -L) and libsusy_hip_U2.so linked
# in their place; oracle/hip_bind.c registers the driver's globals through
# --wrap of readin and compute_Fmunu.  The library is found through a run
# path relative to the binary, so the pair can move together; MKL's directory
# is a run path of the binary only (not of the HIP runtime, which needs the
# system's C++ library).
HIPLIB  := susy_amd/lib
DROPIN  := fermion_op DSq congrad_multi fermion_force gauge_force
LOCALISED := utilities congrad_multi update_h
$(OBJ)/susy_%.local.o: $(OBJ)/susy_%.o
	objcopy $(DROPIN:%=-L %) $< $@
$(OBJ)/hip_bind.o: oracle/hip_bind.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@
OBJS_DROPIN := $(filter-out $(LOCALISED:%=$(OBJ)/susy_%.o),$(OBJS)) \
               $(LOCALISED:%=$(OBJ)/susy_%.local.o)
$(OUT)/susy_hmc_U2_hip: $(OBJS_DROPIN) $(OBJ)/control.o $(OBJ)/hip_bind.o $(HIPLIB)/libsusy_hip_U2.so
	$(CC) -o $@ $(OBJS_DROPIN) $(OBJ)/control.o $(OBJ)/hip_bind.o \
	  -Wl,--wrap=readin -Wl,--wrap=compute_Fmunu \
	  -L$(HIPLIB) -lsusy_hip_U2 $(MKLDIR)/libmkl_rt.so -lm \
	  -Wl,-rpath-link,/usr/lib/x86_64-linux-gnu -Wl,--enable-new-dtags \
	  '-Wl,-rpath,$$ORIGIN/../../susy_amd/lib:$(MKLDIR)'

# ---------------------------------------------------------------------------
# The CPU baseline: the reference at U(4) as an MPI program (Make_mpi's
# com_mpi.c), with oracle/ref_bench.c as its main.  The reference selects the
# gauge group by the NCOL define of include/susy.h:16 and every source reaches
# that header as "../include/susy.h"; the build generates a copy of the header
# with NCOL 4 under $(OUT4)/hdr/include and compiles with -I- so that quoted
# includes are searched in $(OUT4)/hdr/susy first (where ../include/susy.h is
# the generated copy) and then in the reference's own directories.
MPICC  ?= /opt/conda/bin/mpicc
OUT4   := $(OUT)/u4mpi
OBJ4   := $(OUT4)/obj
HDR4   := $(OUT4)/hdr
CFLAGS4 := -O3 -fPIC -fno-strict-aliasing -w -DMPI -DPRECISION=2 \
           -D_FILE_OFFSET_BITS=64 -D_LARGEFILE_SOURCE64 -DPHI_ALGORITHM -DHMC_ALGORITHM \
           -I$(HDR4)/susy -I- -I$(REF)/susy -I$(REF)/generic
GEN4_SRC := $(filter-out com_vanilla,$(GEN_SRC)) com_mpi
OBJS4 := $(SUSY_SRC:%=$(OBJ4)/susy_%.o) $(GEN4_SRC:%=$(OBJ4)/gen_%.o) \
         $(LIB_SRC:%=$(OBJ4)/lib_%.o)

$(HDR4)/include/susy.h: $(REF)/include/susy.h
	mkdir -p $(HDR4)/include $(HDR4)/susy $(OBJ4)
	sed 's/^#define NCOL 2$$/#define NCOL 4/' $< > $@
	grep -q '^#define NCOL 4$$' $@

$(OBJ4)/susy_%.o: $(REF)/susy/%.c $(HDR4)/include/susy.h
	MPICH_CC=$(CC) $(MPICC) $(CFLAGS4) -c $< -o $@
$(OBJ4)/gen_%.o: $(REF)/generic/%.c $(HDR4)/include/susy.h
	MPICH_CC=$(CC) $(MPICC) $(CFLAGS4) -c $< -o $@
$(OBJ4)/lib_%.o: $(REF)/libraries/%.c $(HDR4)/include/susy.h
	MPICH_CC=$(CC) $(MPICC) $(CFLAGS4) -c $< -o $@
$(OBJ4)/ref_harness.o: oracle/ref_harness.c $(HDR4)/include/susy.h
	MPICH_CC=$(CC) $(MPICC) $(CFLAGS4) -c $< -o $@
$(OBJ4)/ref_bench.o: oracle/ref_bench.c $(HDR4)/include/susy.h
	MPICH_CC=$(CC) $(MPICC) $(CFLAGS4) -c $< -o $@

$(OUT)/susy_refbench_U4: $(OBJS4) $(OBJ4)/ref_harness.o $(OBJ4)/ref_bench.o
	MPICH_CC=$(CC) $(MPICC) -o $@ $^ $(LIBS)

clean:
	rm -rf $(OUT)
