# Builds the reference OpenCFD-EC Fortran (test infrastructure only) into oracle/_ref/.
#
# Sources are compiled where they lie under $(REF); nothing from them is copied into the
# repository.  The main-program file opencfd_ec3d_v1.16a.f90 carries module Flow_Var and the
# reference's own `program main`; the driver here replaces the program, so only the Flow_Var
# module text is extracted from that file into a transient scratch directory (removed again
# after the compile).  The extraction also sidesteps the one GBK full-width comment character
# on line 109 of that file that flang rejects.
#
#   make -f oracle/ref.mk            (from the repository root)
#   make -f oracle/ref.mk dropin     the reference host with libocfd_ec.so dropped in through the
#                                    Fortran binding (opencfd-ec_amd/fortran/ocfd_ec.f90); needs
#                                    opencfd-ec_amd/lib/libocfd_ec.so
#
REF      ?= /root/reference
OUT      := oracle/_ref
FC       := /opt/rocm/bin/amdflang
MPIINC   := /opt/conda/include
MPILIB   := -L/opt/conda/lib -lfmpich -lmpich -Wl,-rpath,/opt/conda/lib
FFLAGS   := -O2 -ffp-contract=off -I$(MPIINC)

# every reference source except the one holding `program main` (order: module providers first)
SRCS := sub_modules.f90 sub_interfaces.f90 sub_scheme.f90 sub_flux_split.f90 \
        sub_read_parameter.f90 sub_partation_mpi.f90 sub_convert_inp.f90 sub_IO.f90 \
        sub_geometry.f90 sub_comput_dw.f90 sub_init.f90 sub_Residual.f90 sub_boundary.f90 \
        sub_turbulence_BL.f90 sub_turbulence_SA.f90 sub_turbulence_NewSA.f90 \
        sub_turbulence_SST.f90 sub_time_advance.f90 sub_update_buffer_mpi.f90 sub_Post.f90 \
        sub_time_acceleraction.f90 sub_LU_SGS.f90 sub_debug.f90 sub_filtering.f90 \
        sub_boundary_user.f90 sub_udate_Meshlink_mpi.f90 sub_Finite_Difference1.f90 \
        sub_Finite_Difference2.f90 sub_Post_timeAverage.f90 sub_limitflow.f90

.PHONY: all serial omp dropin clean
all: serial omp

serial: $(OUT)/ref_driver
omp: $(OUT)/ref_driver_omp

define BUILD_VARIANT
$(OUT)/$(1)/.stamp: $(addprefix $(REF)/,$(SRCS)) oracle/ref.mk
	@mkdir -p $(OUT)/$(1)
	@set -e; scratch=$$$$(mktemp -d); \
	sed -n '/^ *module  *Flow_Var/,/^ *end  *module  *Flow_Var/p' $(REF)/opencfd_ec3d_v1.16a.f90 > $$$$scratch/flow_var.f90; \
	cd $(OUT)/$(1); \
	$(FC) $(FFLAGS) $(2) -c $(REF)/sub_modules.f90 -o sub_modules.o; \
	$(FC) $(FFLAGS) $(2) -c $$$$scratch/flow_var.f90 -o flow_var.o; \
	for s in $(filter-out sub_modules.f90,$(SRCS)); do \
	  $(FC) $(FFLAGS) $(2) -c $(REF)/$$$$s -o $$$${s%.f90}.o 2> $$$${s%.f90}.log || { cat $$$${s%.f90}.log; rm -rf $$$$scratch; exit 1; }; \
	done; \
	rm -rf $$$$scratch; touch .stamp

$(OUT)/$(3): $(OUT)/$(1)/.stamp oracle/ref_driver.f90
	cd $(OUT)/$(1) && $(FC) $(FFLAGS) $(2) -c ../../../oracle/ref_driver.f90 -o ref_driver.o && \
	  $(FC) $(2) -o ../$(3) ref_driver.o flow_var.o $(SRCS:.f90=.o) $(MPILIB)
endef

$(eval $(call BUILD_VARIANT,obj_serial,,ref_driver))
$(eval $(call BUILD_VARIANT,obj_omp,-fopenmp,ref_driver_omp))

HIPLIB := opencfd-ec_amd/lib
dropin: $(OUT)/dropin_driver
$(OUT)/dropin_driver: $(OUT)/obj_serial/.stamp oracle/dropin_driver.f90 opencfd-ec_amd/fortran/ocfd_ec.f90 $(HIPLIB)/libocfd_ec.so
	cd $(OUT)/obj_serial && $(FC) $(FFLAGS) -c ../../../opencfd-ec_amd/fortran/ocfd_ec.f90 -o ocfd_ec_binding.o && \
	  $(FC) $(FFLAGS) -c ../../../oracle/dropin_driver.f90 -o dropin_driver.o && \
	  $(FC) -o ../dropin_driver dropin_driver.o ocfd_ec_binding.o flow_var.o $(SRCS:.f90=.o) $(MPILIB) \
	    -L../../../$(HIPLIB) -locfd_ec '-Wl,-rpath,$$ORIGIN/../../$(HIPLIB)'

clean:
	rm -rf $(OUT)
