# Builds the reference's own Fortran modules, unmodified and in place under
# /root/reference, into oracle/_ref/ (objects + .mod files), then links the
# fixture dumper refdump.F90 (our code); `make -f oracle/ref.mk mixture`
# also links the reference MD driver Mixture.
# Builds the reference from its own sources on the host that holds them; nothing
# built here is loaded by the GPU tests or the benchmark.
#
#   make -f oracle/ref.mk            (from the repository root)
#
REF     ?= /root/reference
OUT     ?= oracle/_ref
REFFC   ?= flang
# -O2 on baseline x86-64: no FMA instructions, so no contraction.
FFLAGS  ?= -cpp -DSTDERR=0 -O2
FFLAGS  += -J$(OUT)

# Module dependency order of the objects the Mixture link line names, plus
# PairInfo/Neighbor/MonteCarloModule (derived from the reference's .depend).
OBJS = Umbrella Error PhysConst CommonModule RandomModule Graph Replica MDReplica \
  MatrixModule VectorModule Monitor QuatModule SiteModule MolModule BoxModule \
  InteractionModule CutoffModule NoseModule AverageModule TimeModule ExternalModule \
  AndersenModule GridModule BookModule MonteCarloModule StdInteraction Argon FlexModule \
  RattleModule Stack TIP4PConst OPLSMeOHModule RigidModule NvdEModule SJBenzene \
  TIP4PModule Triplet SWSilicon ST2Module Interaction2 FlexSetup EPM2CO2Module \
  SPCEModule FlatFallerTHFModule SSPKMethaneModule CJTHFModule TIP5PModule PiWater \
  Dumbbell RigidSetup PairInfo Neighbor

OBJFILES = $(addprefix $(OUT)/,$(addsuffix .o,$(OBJS))) $(OUT)/mtrnd.o

all: $(OUT)/refdump $(OUT)/refneigh

# The reference MD driver (about 300 MB of static arrays); built only on
# demand by tests/golden/make_golden.py, which removes it afterwards.
mixture: $(OUT)/Mixture

# The reference MC program mc2 (about 280 MB of static arrays; its link line
# is Makefile.rules:32); built on demand by tests/golden/make_mc2.py, which
# removes it afterwards.
MC2OBJS = Tetrahedrality Set DistribModule Triangle q6 Q6Local mc2util
mc2: $(OUT)/mc2

$(OUT):
	mkdir -p $(OUT)

# Serial build: each object needs the .mod files of those before it.
$(OUT)/objs.stamp: | $(OUT)
	set -e; for o in $(OBJS); do \
	  $(REFFC) $(FFLAGS) -c $(REF)/$$o.F90 -o $(OUT)/$$o.o; done
	$(REFFC) -O2 -c $(REF)/mtrnd.f -o $(OUT)/mtrnd.o
	touch $@

# The reference's C helpers (process transport, timer) are needed only at
# link time: MDReplica references them.
COBJS = $(OUT)/FakeMPI.o $(OUT)/FakeMPIF.o $(OUT)/Timer.o
$(OUT)/%.o: $(REF)/%.c | $(OUT)
	gcc -O2 -c $< -o $@

$(OUT)/refdump: oracle/refdump.F90 $(OUT)/objs.stamp $(COBJS)
	$(REFFC) $(FFLAGS) oracle/refdump.F90 $(OBJFILES) $(COBJS) -o $@

$(OUT)/refneigh: oracle/refneigh.F90 $(OUT)/objs.stamp $(COBJS)
	$(REFFC) $(FFLAGS) oracle/refneigh.F90 $(OBJFILES) $(COBJS) -o $@

$(OUT)/mc2.stamp: $(OUT)/objs.stamp
	set -e; for o in $(MC2OBJS); do \
	  $(REFFC) $(FFLAGS) -c $(REF)/$$o.F90 -o $(OUT)/$$o.o; done
	touch $@

# the order-parameter dumper (needs mc2's modules); built on demand by
# tests/golden/make_order.py, which removes it afterwards
$(OUT)/refop: oracle/refop.F90 $(OUT)/mc2.stamp $(COBJS)
	$(REFFC) $(FFLAGS) oracle/refop.F90 $(OBJFILES) $(addprefix $(OUT)/,$(addsuffix .o,$(MC2OBJS))) \
	  $(COBJS) -o $@

$(OUT)/mc2: $(OUT)/mc2.stamp $(COBJS) $(OUT)/om.o
	$(REFFC) $(FFLAGS) -c $(REF)/mc2.F90 -o $(OUT)/mc2.o
	$(REFFC) $(FFLAGS) $(OUT)/mc2.o $(OBJFILES) $(addprefix $(OUT)/,$(addsuffix .o,$(MC2OBJS))) \
	  $(OUT)/om.o $(COBJS) -o $@
	rm -f $(OUT)/mc2.o

# The reference MD driver itself.
$(OUT)/Mixture: $(OUT)/objs.stamp $(COBJS)
	$(REFFC) $(FFLAGS) -c $(REF)/Mixture.F90 -o $(OUT)/Mixture.o
	$(REFFC) $(FFLAGS) $(OUT)/Mixture.o $(OBJFILES) $(COBJS) -o $@
	rm -f $(OUT)/Mixture.o

clean:
	rm -rf $(OUT)

.PHONY: all clean mixture
