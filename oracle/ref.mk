# Builds the reference CHOLMOD (CPU path: SuiteSparse_config, AMD, CAMD, COLAMD,
# CCOLAMD, CHOLMOD with its bundled METIS) straight from the sources under
# /root/reference into oracle/_ref/libssref.so, plus the small driver
# oracle/ref_driver.c that the tests and the CPU-baseline leg call through
# ctypes.  Nothing is copied out of /root/reference; objects and the library
# land in oracle/_ref/ only (git-ignored, carried to the GPU host as built).
#
# BLAS/LAPACK: MKL from the image (/opt/conda/lib/libmkl_rt.so, LP64).  The
# driver forces MKL_THREADING_LAYER=GNU before the first BLAS call: MKL's
# Intel-OpenMP layer mixed with CHOLMOD's libgomp gives wrong factors.
#
#   make -f oracle/ref.mk -j8          (from the repository root)

REF      ?= /root/reference
OUT      := oracle/_ref
OBJ      := $(OUT)/obj
MKLDIR   ?= /opt/conda/lib
CC       := gcc
CFLAGS   := -O3 -fPIC -fopenmp -DNDEBUG -std=gnu11 -w -DBLAS_UNDERSCORE -DBLAS32
INC      := -I$(REF)/SuiteSparse_config -I$(REF)/AMD/Include -I$(REF)/CAMD/Include \
            -I$(REF)/COLAMD/Include -I$(REF)/CCOLAMD/Include -I$(REF)/CHOLMOD/Include \
            -I$(REF)/CHOLMOD -I$(REF)/CHOLMOD/Check -I$(REF)/CHOLMOD/Cholesky \
            -I$(REF)/CHOLMOD/Utility -I$(REF)/CHOLMOD/MatrixOps -I$(REF)/CHOLMOD/Modify \
            -I$(REF)/CHOLMOD/Partition -I$(REF)/CHOLMOD/Supernodal \
            -I$(REF)/CHOLMOD/SuiteSparse_metis/GKlib -I$(REF)/CHOLMOD/SuiteSparse_metis/libmetis \
            -I$(REF)/CHOLMOD/SuiteSparse_metis/include

SRCS := $(REF)/SuiteSparse_config/SuiteSparse_config.c \
        $(wildcard $(REF)/AMD/Source/amd_*.c) \
        $(wildcard $(REF)/CAMD/Source/camd_*.c) \
        $(wildcard $(REF)/COLAMD/Source/colamd*.c) \
        $(wildcard $(REF)/CCOLAMD/Source/ccolamd*.c) \
        $(wildcard $(REF)/CHOLMOD/Check/cholmod_*.c) \
        $(wildcard $(REF)/CHOLMOD/Cholesky/cholmod_*.c) \
        $(wildcard $(REF)/CHOLMOD/Utility/cholmod_*.c) \
        $(wildcard $(REF)/CHOLMOD/MatrixOps/cholmod_*.c) \
        $(wildcard $(REF)/CHOLMOD/Modify/cholmod_*.c) \
        $(wildcard $(REF)/CHOLMOD/Partition/cholmod_*.c) \
        $(wildcard $(REF)/CHOLMOD/Supernodal/cholmod_*.c)

# one object per source, named by its path below $(REF)
objname = $(OBJ)/$(subst /,__,$(patsubst $(REF)/%.c,%,$(1))).o
OBJS := $(foreach s,$(SRCS),$(call objname,$(s)))

all: $(OUT)/libssref.so

define compile_rule
$(call objname,$(1)): $(1) | $(OBJ)
	$(CC) $(CFLAGS) $(INC) -c $$< -o $$@
endef
$(foreach s,$(SRCS),$(eval $(call compile_rule,$(s))))

$(OBJ)/ref_driver.o: oracle/ref_driver.c | $(OBJ)
	$(CC) $(CFLAGS) $(INC) -c $< -o $@

$(OUT)/libssref.so: $(OBJS) $(OBJ)/ref_driver.o
	$(CC) -shared -fopenmp -o $@ $^ -L$(MKLDIR) -Wl,-rpath,$(MKLDIR) -lmkl_rt -lm

$(OBJ):
	mkdir -p $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
