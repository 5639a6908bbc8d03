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

all: $(OUT)/libsusyref_U2.so $(OUT)/susy_hmc_U2

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

clean:
	rm -rf $(OUT)
