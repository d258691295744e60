

Bin/rtm.x:   lin7@cees-gpu   Wed Jul 22 09:18:26 2015
		n1=300
		o1=0
		d1=4
		n2=300
		o2=0
		d2=4
		sets next: in="/data/lin7/drho-layer.H@"
		data_format="xdr_float"


Cp:   lin7@cees-gpu   Wed Jul 22 09:19:44 2015
		sets next: in="/data/lin7/drho-layer-check.H@"
		copy_intag="Dat/drho-layer.H"
		data_format="xdr_float"
