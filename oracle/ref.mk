# Builds oracle/_ref/libffo_ref.so from the reference's own dependency-free
# modules (compiled where they lie under /root/reference) plus
# oracle/ref_bindings.f90.  This container only; the reference's other
# modules need fortsparse / fortnum, which are absent, and are not built.
#   make -C oracle -f ref.mk
REF    ?= /root/reference
FLANG  ?= /opt/rocm/lib/llvm/bin/flang
OUTDIR := _ref
SRCS   := $(REF)/src/utils/kinds.f90 \
          $(REF)/src/elements/tetra_lagrange_arbitrary_order.f90 \
          $(REF)/src/elements/tetra_lagrange_global_dof_map.f90 \
          $(REF)/src/elements/tetra_edge_dof_map.f90 \
          $(REF)/src/elements/tetra_nedelec_first_order.f90 \
          $(REF)/src/generated/fortfem_tetra_nedelec_coefficients.f90 \
          $(REF)/src/generated/fortfem_tetra_nedelec_candidates_order_1.f90 \
          $(REF)/src/generated/fortfem_tetra_nedelec_candidates_order_2.f90 \
          $(REF)/src/generated/fortfem_tetra_nedelec_candidates_order_3.f90 \
          $(REF)/src/generated/fortfem_tetra_nedelec_candidates_order_4.f90 \
          $(REF)/src/generated/fortfem_tetra_face_moment_transforms.f90 \
          $(REF)/src/generated/fortfem_tetra_h1_oracle.f90 \
          ref_bindings.f90

$(OUTDIR)/libffo_ref.so: $(SRCS)
	@mkdir -p $(OUTDIR)/mod
	$(FLANG) -O1 -fPIC -shared -module-dir $(OUTDIR)/mod -o $@.tmp $(SRCS) && mv $@.tmp $@

clean:
	rm -rf $(OUTDIR)

.PHONY: clean
